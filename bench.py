#!/usr/bin/env python3
"""Throughput of the batched 2D FDTD hot path on one MI355X.

Workload (one bench "step" = one full simulation of one batch): 2D TMz,
600 x 600 cells, dx = 0.5 mm, Courant 0.5, 15-cell CPML (the stepped box is the
571 x 571 cavity the reference's inert CPML encloses), 16 TX = 16 RX on a
200-cell ring, 800 time steps, single-pole Debye head phantom (seed 0), then
the running DFT of all 256 traces at 301 frequencies over 0.5-3 GHz into the
(301, 16, 16) S-matrix.  Inputs are resident on the device before timing.

Prints one JSON line.  Its `roofline` is for the dominant kernel: `traffic`
is the HBM bytes one launch (K steps of the whole batch) moves by rocprofv3
counter (the committed summary under profiles/; with the TX in groups on
separate streams a launch is `dispatches_per_launch` dispatches, each counted),
`achieved` and `frac` that over the launch's mean HIP-event time;
`one_dispatch` the same K steps as a single dispatch over the whole batch;
`achieved_algorithmic` / `frac_algorithmic` count every stepped cell's fields
and ADE current read and written once per TX plus its maps once, times the
steps per launch -- the temporal blocking moves less than that, so that figure
exceeds the peak.  The `adjoint` leg carries the reversible adjoint kernel's
roofline the same way (`roofline_adjoint`).
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import platform
import sys
import time

# a fatal signal (an abort from a heap check, a segmentation fault) prints the
# Python stack of every thread to stderr before the process ends
faulthandler.enable()

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ceep-v1_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

try:
    # torch first: libceep_fdtd.so then binds to torch's HIP runtime (same
    # SONAME) instead of loading a second one, and torch.cuda stays usable
    import torch  # noqa: E402,F401
except Exception:  # pragma: no cover - torch is optional for the bench
    torch = None

NX = NY = 600
DX = 0.5e-3
N_CPML = 15
T = 800
B = 16
COURANT = 0.5
FREQS = np.linspace(0.5e9, 3e9, 301)
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-step: ez, hx, hy, J read + written once (32 B per
# TX) and the shared ca, cb, k1, k2 maps read once per launch (16 B per cell)
BYTES_PER_CELL_PER_TX = 32
BYTES_PER_CELL_SHARED = 16
# 2D reversible adjoint, per reversed cell-step and TX: the forward's E, Hx, Hy
# and the adjoint's lambda_e, mu_x, mu_y read + written once
BYTES_ADJ_PER_CELL_PER_TX = 48
BYTES3D_PER_CELL_PER_TX = 120   # 3D Debye + Lorentz: 15 state floats read + written
BYTES3D_SHARED = 28             # ca, cb, two Debye and three Lorentz coefficient maps


def _workload():
    import ceep_amd
    from ceep_amd.constants import cfl_timestep_2d
    from ceep_amd.waveforms import modulated_gaussian

    dt = cfl_timestep_2d(DX, DX, COURANT)
    pos = ceep_amd.phantoms.ring_positions((300, 300), 200, B)
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    ph = ceep_amd.phantoms.DebyeHeadPhantom2D(NX, NY, 0)
    s = ceep_amd.BatchedFDTD2D(nx=NX, ny=NY, dx=DX, total_steps=T, cpml_thickness=N_CPML,
                               source_positions=pos, probe_positions=pos, frequency=2e9,
                               courant=COURANT, waveform=w, cpml_R=1e-6)
    s.set_phantom(ph)
    s._build()
    return s, ph, pos, w, dt


def _progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _device_sync():
    # every library call already returns after its stream has drained; the
    # device-wide synchronize is the contract's barrier on top of that
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()
        return "torch.cuda.synchronize"
    return "library-synchronous"


def _one_core(fn):
    """Run fn pinned to one host core (NumPy's elementwise kernels are
    single-threaded; the pin keeps BLAS and torch threads out of the timing)."""
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        return fn()
    finally:
        os.sched_setaffinity(0, old)


def _cpu_baseline(ph, pos, w, dt, steps):
    """The oracle's FP64 NumPy restatements of the reference's CPU path on one
    pinned host core, first `steps` steps of the workload: "ref-cost" keeps the
    reference's per-row CPML loops (fdtd_2d_batched.py:272-406, its CPU cost
    profile) and is `value`; "vectorised" is the run_cpu-style full-array form
    (:540-611) on the stepped box, with the same Debye ADE."""
    from oracle import ade, fdtd2d

    m = ph.get_debye_maps()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)]

    def ref_cost(n):
        fdtd2d.run_literal(NX, NY, DX, dt, N_CPML, pos, pos, w[:n], m["eps_inf"], m["sigma"],
                           poles=poles, ref_cost=True)

    def vectorised(n):
        fdtd2d.run_cavity(NX, NY, DX, dt, N_CPML, pos, pos, w[:n], m["eps_inf"], m["sigma"],
                          poles=poles)

    def timed(fn):
        fn(2)                                   # warm-up
        t0 = time.perf_counter()
        fn(steps)
        return time.perf_counter() - t0

    el = _one_core(lambda: timed(ref_cost))
    el_v = _one_core(lambda: timed(vectorised))
    box = (NX - 2 * N_CPML + 1) ** 2
    host = f"{_cpu_model()}, os.cpu_count()={os.cpu_count()}"
    return {
        "value": NX * NY * B * steps / el / 1e9,
        "unit": "GCell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"first {steps} of {T} steps of this workload (full 600x600 grid, B=16, "
                   f"Debye) through the oracle's FP64 NumPy restatement with the reference's "
                   f"per-row CPML loops, pinned to one core, {el:.1f} s; host {host}"),
        "vectorised": {
            "value": round(box * B * steps / el_v / 1e9, 5),
            "unit": "GCell-steps/s (stepped box)",
            "value_full_grid_convention": round(NX * NY * B * steps / el_v / 1e9, 5),
            "cores": 1,
            "sample": (f"first {steps} of {T} steps, the run_cpu-style vectorised FP64 NumPy "
                       f"form on the 571x571 box with the Debye ADE, pinned to one core, "
                       f"{el_v:.1f} s"),
        },
    }


def _cpu_openmp(ph, pos, w, dt, steps=100):
    """The oracle's C restatement (FP64, OpenMP over every host core) on the
    first `steps` steps of the same workload: a stronger CPU reference point
    than the reference's own single-core NumPy path."""
    from oracle import ade, c_ref

    m = ph.get_debye_maps()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)]
    c_ref.run_cavity2d(NX, NY, DX, dt, N_CPML, pos, pos, w[:5], m["eps_inf"], m["sigma"],
                       poles=poles)
    t0 = time.perf_counter()
    c_ref.run_cavity2d(NX, NY, DX, dt, N_CPML, pos, pos, w[:steps], m["eps_inf"], m["sigma"],
                       poles=poles)
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round((NX - 2 * N_CPML + 1) ** 2 * B * steps / el / 1e9, 4),
            "unit": "GCell-steps/s (stepped box)", "cores": threads, "kind": "port",
            "sample": (f"first {steps} of {T} steps, FP64 C restatement, {threads} OpenMP threads, "
                       f"{el:.2f} s; host {_cpu_model()}")}


def _gradient_workload():
    """The 2D gradient workload: 600x600, 16 TX/RX ring r=200, 800 steps, seeded
    lossy disk (eps_r 10-60, sigma 0.2-1.5, radius 30-90 cells; seed 0); the
    misfit's target is the disk of seed 1, S at 16 frequencies in 1-2.5 GHz."""
    import ceep_amd
    from ceep_amd.constants import cfl_timestep_2d
    from ceep_amd.waveforms import modulated_gaussian

    dt = cfl_timestep_2d(DX, DX, COURANT)
    pos = ceep_amd.phantoms.ring_positions((300, 300), 200, B)
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    freqs = np.linspace(1e9, 2.5e9, 16)
    kw = dict(r=(30.0, 90.0), max_offset=120.0)
    eps, sig = ceep_amd.phantoms.DiskPhantom2D(NX, NY, 0, **kw).get_eps_map(2e9)
    eps_t, sig_t = ceep_amd.phantoms.DiskPhantom2D(NX, NY, 1, **kw).get_eps_map(2e9)

    def solver(e, sg, tuning=None):
        s = ceep_amd.BatchedFDTD2D(nx=NX, ny=NY, dx=DX, total_steps=T, cpml_thickness=N_CPML,
                                   source_positions=pos, probe_positions=pos, frequency=2e9,
                                   courant=COURANT, waveform=w, cpml_R=1e-6)
        s._eps_r[:] = e
        s._sigma_e[:] = sg
        if tuning is not None:
            from ceep_amd.autotune import tuning_for
            s.tuning = {**tuning_for(2, (NX, NY), N_CPML, B, False, False), **tuning}
        return s

    s = solver(eps, sig)
    s._build()
    return s, eps, sig, freqs, (eps_t, sig_t), solver, dt


def _fp64_fd(solver, dt, freqs, eps, sig, eps_t, sig_t, ge, gs, n_pix=4, seed=6, h=1e-4):
    """Worst relative deviation of (ge, gs) from FP64 central differences
    (h = 1e-4) of L = sum |S - S_target|^2 on n_pix seeded pixels among the 400
    largest |gradient| cells (half eps_r, half sigma; the pixel choice of the
    600^2 finite-difference test).  The losses come from the engine's FP64
    forward kernels (bitwise the FP64 oracle) and an FP64 DFT on the host."""
    s64 = solver(eps, sig, {"fp64_forward": 1})
    s64._build()
    ph = np.exp(-2j * np.pi * freqs[:, None] * (np.arange(T) * dt)[None, :]) * dt

    def S_of(e, sg):
        s64.update_materials(e, sg)
        return np.einsum("ft,bpt->fpb", ph, s64._handle.run_checkpointed(T))

    St = S_of(eps_t, sig_t)

    def loss(e, sg):
        return float(np.sum(np.abs(S_of(e, sg) - St) ** 2))

    rng = np.random.default_rng(seed)
    worst = 0.0
    L64 = loss(eps, sig)
    for grad, which in ((ge, "eps"), (gs, "sigma")):
        top = np.argsort(np.abs(grad).ravel())[-400:]
        for flat in rng.choice(top, size=n_pix // 2, replace=False):
            a, b = np.unravel_index(flat, grad.shape)
            base = eps if which == "eps" else sig
            p, m = base.copy(), base.copy()
            p[a, b] += h
            m[a, b] -= h
            fd = ((loss(p, sig) - loss(m, sig)) if which == "eps"
                  else (loss(eps, p) - loss(eps, m))) / (2 * h)
            worst = max(worst, abs(grad[a, b] - fd) / abs(fd))
    s64.close()
    return worst, L64, St


def _adjoint_ratio(every=50):
    """fwd+adjoint / fwd on the gradient workload (see _gradient_workload):
    loss = sum |S - S_target|^2 at 16 frequencies in 1-2.5 GHz against the
    target disk, checkpoints every `every` steps, FP32 difference form (the
    engine steps E(eps) - E(target); the variant the 600^2 finite-difference
    test checks).  The target's own run is taken once per target and timed
    apart."""
    from ceep_amd import adjoint

    s, eps, sig, freqs, (eps_t, sig_t), solver, dt = _gradient_workload()
    h = s._handle
    for _ in range(2):
        h.reset()
        h.run(copy=False)
        h.dft(freqs)
    _device_sync()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        h.reset()
        h.run(copy=False)
        h.dft(freqs)
    _device_sync()
    t_fwd = (time.perf_counter() - t0) / reps

    def timed(fn):
        fn()      # warm (allocations)
        _device_sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        _device_sync()
        return (time.perf_counter() - t0) / reps, out

    # the target's run (once per target): the first residual run minus a later one
    adjoint.set_target(s, eps_t, sig_t)
    _device_sync()
    t0 = time.perf_counter()
    h.run_residual(every)
    _device_sync()
    t_first = time.perf_counter() - t0
    t_res, _ = timed(lambda: h.run_residual(every))
    # the reversible adjoint on its own, for its roofline: a residual run and
    # the adjoint of a fixed dL/dtrace, minus the residual run alone
    res, _ = h.run_residual(every)
    r = torch.from_numpy(res).requires_grad_(True)
    adjoint.s_norm2(r, freqs, dt).backward()
    g = r.grad.numpy().copy()
    t_ra, _ = timed(lambda: (h.run_residual(every), h.adjoint(g)))
    box = h.box_info()
    cells = int(box[0] * box[1])
    n_steps, n_tx = s.total_steps, s.batch
    k_adj, bpw_adj = 6, 4
    launches = -(-n_steps // every) * -(-every // k_adj)
    chunks = -(-n_tx // bpw_adj)
    streams = (s.tuning or {}).get("tx_streams", 2)
    # algorithmic bytes per launch of K reversed steps: the forward's E, Hx, Hy
    # and the adjoint's lambda_e, mu_x, mu_y read and written once per cell,
    # TX and step; the two gradient sums per TX chunk and the two maps once
    bytes_adj = cells * (n_tx * BYTES_ADJ_PER_CELL_PER_TX * k_adj + chunks * 16 + 8)
    roof_adj = _roofline(f"adj2d_rev_mw_kernel<8, {k_adj}, {bpw_adj}, false>", k_adj, bytes_adj,
                         (t_ra - t_res) * 1e3 / launches, "gradient2d_disk_600x600x16_counters.json",
                         min(streams, chunks), section="two_stream_gradient_run")
    roof_adj["ms_adjoint_call"] = round((t_ra - t_res) * 1e3, 3)
    t_grad, (loss, ge, gs) = timed(lambda: adjoint.gradient(s, eps, sig, freqs, every=every,
                                                            target=(eps_t, sig_t)))
    t_grad_ag, _ = timed(lambda: adjoint.gradient(s, eps, sig, freqs, every=every,
                                                  target=(eps_t, sig_t), fused=False))
    worst, L64, St64 = _fp64_fd(solver, dt, freqs, eps, sig, eps_t, sig_t, ge, gs)
    t_grad64, _ = timed(lambda: adjoint.gradient(s, eps, sig, freqs, St64, every=every,
                                                 fp64_forward=True))
    s.close()
    return {"ratio": round(t_grad / t_fwd, 3), "ms_forward": round(t_fwd * 1e3, 3),
            "ms_forward_plus_adjoint": round(t_grad * 1e3, 3), "checkpoint_every": every,
            "fd_worst_rel_dev": float(f"{worst:.3e}"),
            "fd_pixels": 4,
            "loss": float(f"{loss:.6e}"), "loss_fp64": float(f"{L64:.6e}"),
            "ms_target_run_once": round((t_first - t_res) * 1e3, 3),
            "ratio_fp64_forward": round(t_grad64 / t_fwd, 3),
            "ratio_autograd_path": round(t_grad_ag / t_fwd, 3),
            "roofline_adjoint": roof_adj,
            "note": ("ratio: the FP32 difference form (forward of E(eps) - E(target) with "
                     "checkpoints every 50 steps, reversible adjoint, chain rule on the device) "
                     "in one library call, S-matrix, loss and dL/dtrace on the device; "
                     "ratio_autograd_path: the same gradient through torch.autograd with the "
                     "misfit on the host (the path for other losses); "
                     "fd_worst_rel_dev: that gradient against FP64 central differences "
                     "(h = 1e-4, 2 eps_r + 2 sigma pixels) from the engine's FP64 forward; "
                     "ms_target_run_once: the target's own run with its E kept at every step, "
                     "taken once per target (like S_target itself) and outside the ratio; "
                     "ratio_fp64_forward: the other accurate variant, FP64 forward fields "
                     "against an FP64 S_target; roofline_adjoint: the reversible adjoint "
                     "kernel, timed as one adjoint call (the residual run, then the adjoint "
                     "of a fixed dL/dtrace, minus the residual run; the call's chain rule "
                     "and copies included) over its launches"),
            "workload": ("2D 600x600, 16 TX/RX ring r=200, 800 steps, lossy disk phantom; "
                         "L = sum |S - S_target|^2 at 16 freqs in 1-2.5 GHz against the disk "
                         "of seed 1; dL/deps_r, dL/dsigma")}


def _bench_3d(steps_timed=2, T3=1000, tuning=None):
    """3D 200^3, dx = 1 mm, 10-cell CPML, 16 z-directed sources = probes on a
    70-cell ring at z = 100, Debye + Lorentz head (seed 0), 1000 steps."""
    import ceep_amd
    from ceep_amd.constants import cfl_timestep_3d
    from ceep_amd.waveforms import modulated_gaussian

    n3, dx3 = 200, 1e-3
    dt3 = cfl_timestep_3d(dx3, dx3, dx3, 0.5) * 0.99
    ring = ceep_amd.phantoms.ring_positions((100, 100), 70, 16)
    pos = [(x, y, 100) for x, y in ring]
    w = modulated_gaussian(T3, dt3, 2e9, 1.5e9)
    ph = ceep_amd.phantoms.DebyeHeadPhantom3D((n3, n3, n3), 0)
    s = ceep_amd.BatchedFDTD3D(nx=n3, ny=n3, nz=n3, dx=dx3, total_steps=T3, cpml_thickness=10,
                               source_positions=pos, probe_positions=pos, frequency=2e9,
                               waveform=w)
    s.set_phantom(ph)
    if tuning is not None:
        s.tuning = tuning
    s._build()
    h = s._handle
    box = h.box_info()
    cells = int(box[0] * box[1] * box[2])
    h.time_runs(1)
    ms = h.time_runs(steps_timed)
    s.close()
    form = (tuning or {}).get("fused_3d", 4)
    k3 = (tuning or {}).get("temporal_depth_3d", 2) if form == 4 else 1
    launches = -(-T3 // k3)
    # algorithmic bytes per step: the 15-float state (E, H, Debye J, Lorentz J
    # and J one step older) read and written once per cell and TX, the 7 maps
    # once per cell; a K-step launch is credited K steps
    bytes_launch = cells * (16 * BYTES3D_PER_CELL_PER_TX + BYTES3D_SHARED) * k3
    kernel = f"step3d_tb_kernel<{k3}, 3, false, false>" if form == 4 else "h3d_kernel + e3d_kernel"
    return {"value": round(cells * 16 * T3 / (ms * 1e-3) / 1e9, 3), "unit": "GCell-steps/s",
            "ms_per_run": round(ms, 2), "cells_counted": f"stepped box {box[0]}x{box[1]}x{box[2]}",
            "value_full_grid_convention": round(n3 ** 3 * 16 * T3 / (ms * 1e-3) / 1e9, 3),
            "roofline": _roofline(kernel, k3, bytes_launch, ms / launches,
                                  "step3d_debye_lorentz_200x200x200x16_traffic.json"),
            "workload": ("3D 200^3 dx=1mm, 10-cell CPML, 16 Ez TX=RX on r=70 ring at z=100, "
                         f"1000 steps, Debye+Lorentz head seed 0 ({kernel}, {k3} steps per launch)")}


def _uneven_solvers(n_problems=64, seed=4):
    """The uneven batch's 64 seeded Debye heads (built solvers) and their grid sizes."""
    import ceep_amd
    from ceep_amd.constants import cfl_timestep_2d
    from ceep_amd.waveforms import modulated_gaussian

    rng = np.random.default_rng(seed)
    dt = cfl_timestep_2d(DX, DX, COURANT)
    solvers, grids = [], []
    for i in range(n_problems):
        g = int(rng.choice([256, 384, 512, 600]))
        steps = int(rng.integers(600, 1201))
        ncp = int(rng.integers(12, 16))
        pos = ceep_amd.phantoms.ring_positions((g / 2, g / 2), 200.0 * g / 600.0, B)
        w = modulated_gaussian(steps, dt, 2e9, 1.5e9)
        s = ceep_amd.BatchedFDTD2D(nx=g, ny=g, dx=DX, total_steps=steps, cpml_thickness=ncp,
                                   source_positions=pos, probe_positions=pos, frequency=2e9,
                                   courant=COURANT, waveform=w, cpml_R=1e-6)
        s.set_phantom(ceep_amd.phantoms.DebyeHeadPhantom2D(g, g, 1000 + i))
        s._build()
        solvers.append(s)
        grids.append(g)
    return solvers, grids


def _bench_uneven(n_problems=64, seed=4, autotune=True, single=False, persistent=False):
    """64 seeded Debye heads (the head generator above scaled to each grid),
    grids drawn from {256, 384, 512, 600}^2, steps U[600, 1200], 16 TX = RX on a
    ring of radius 200/600 of the grid, CPML 12-15 cells, each followed by its
    301-frequency DFT.  The batch runs as one captured graph, timed in both
    forms -- one persistent launch (four-wavefront shapes) and one launch per
    round -- with the launch shape from the coordinate autotuner; `value` is
    the faster.  The comparison runs the 64 problems one after another with
    the one-step kernel.  `single`: one batch in the persistent form (or per
    round with `persistent` False), nothing timed (for rocprofv3 --pmc)."""
    from ceep_amd.autotune import autotune_uneven, autotune_uneven_persistent
    from ceep_amd.uneven import UnevenBatch

    solvers, grids = _uneven_solvers(n_problems, seed)
    # K, window rows, TX planes per window (per item), wavefronts: the tuners' picks
    shape, log = (8, 64, 4, 8), []
    shape_p, log_p = (6, 48, 2, 4), []
    if autotune:
        shape, log = autotune_uneven(solvers)
        shape_p, log_p = autotune_uneven_persistent(solvers)
    if single:                       # one batch, nothing timed (for rocprofv3 --pmc)
        ub = UnevenBatch(solvers, *(shape_p if persistent else shape))
        ub._multi.run(UnevenBatch.PERSISTENT if persistent else UnevenBatch.PER_ROUND)
        _device_sync()
        ub.close()
        return None
    forms = {"per_round": (shape, UnevenBatch.PER_ROUND), "persistent": (shape_p, UnevenBatch.PERSISTENT)}

    def batched(ub, mode):
        ub._multi.run(mode)
        for s in solvers:
            s._handle.dft(FREQS)

    def sequential():
        for s in solvers:
            s._handle.reset()
            s._handle.run(copy=False)
            s._handle.dft(FREQS)

    t_mode, graph_mode, infos = {}, {}, {}
    for name, (shp, mode) in forms.items():
        ub = UnevenBatch(solvers, *shp)
        batched(ub, mode)
        _device_sync()
        t0 = time.perf_counter()
        batched(ub, mode)
        _device_sync()
        t_mode[name] = time.perf_counter() - t0
        graph_mode[name] = ub.time(2, mode)
        infos[name] = ub.info()
        cell_steps = ub.cell_steps
        ub.close()
    full_cells = sum(g * g * B * s.total_steps for g, s in zip(grids, solvers))
    best = min(t_mode, key=t_mode.get)
    t_b, ms_graph = t_mode[best], graph_mode[best]
    info = infos["per_round"]
    for s in solvers:
        s._handle.apply_tuning({"temporal_depth": 0, "batch_per_thread": 4, "graph": 1})
    sequential()
    _device_sync()
    t0 = time.perf_counter()
    sequential()
    _device_sync()
    t_s = time.perf_counter() - t0
    return {"value": round(cell_steps / t_b / 1e9, 3), "unit": "GCell-steps/s",
            "ms_per_batch": round(t_b * 1e3, 2), "ms_graph_only": round(ms_graph, 2),
            "form": best,
            "ms_per_batch_by_form": {k: round(v * 1e3, 2) for k, v in t_mode.items()},
            "ms_graph_only_by_form": {k: round(v, 2) for k, v in graph_mode.items()},
            "value_full_grid_convention": round(full_cells / t_b / 1e9, 3),
            "sequential_one_step_value": round(cell_steps / t_s / 1e9, 3),
            "sequential_one_step_ms": round(t_s * 1e3, 2),
            "speedup_vs_sequential": round(t_s / t_b, 3),
            "launch_shape": {"temporal_depth": shape[0], "rows_per_wave": shape[1],
                             "batch_per_wave": shape[2], "wavefronts_per_window": shape[3]},
            "launch_shape_persistent": {"temporal_depth": shape_p[0], "rows_per_wave": shape_p[1],
                                        "batch_per_item": shape_p[2]},
            "autotune_log_ms": [[round(ms, 2), list(sh)] for ms, sh in sorted(log)],
            "autotune_log_persistent_ms": [[round(ms, 2), list(sh)] for ms, sh in sorted(log_p)],
            "work_items": info["items"], "segments": info["segments"],
            "persistent_items": infos["persistent"]["persistent_items"],
            "persistent_workgroups": infos["persistent"]["persistent_workgroups"],
            "cell_steps": cell_steps,
            "grids": {str(g): grids.count(g) for g in sorted(set(grids))},
            "workload": (f"{n_problems} seeded Debye heads, grids from 256/384/512/600 squared, "
                         "steps U[600,1200], 16 TX/RX, CPML 12-15, 301-frequency DFT each; "
                         "one hipGraph (one persistent launch, or one launch per round) over all "
                         "problems vs the problems run one after another with the one-step kernel")}


def _bench_3d_gradient(T3=1000, every=100, single=False, tuning=None):
    """dL/d eps_r and dL/d sigma over the whole 200^3 volume of the 3D
    workload (Debye + Lorentz head seed 0, 16 TX, 1000 steps), loss = sum
    |S - S_target|^2 over the 101-frequency S-matrix against head seed 1,
    checkpoints every 100 steps with recomputation.  Ratio = gradient wall
    time over the forward's (reset, run, DFT).  With `single`: one gradient
    call and nothing timed (for rocprofv3 --pmc).  `tuning`: launch settings
    on top of the defaults, for both solvers."""
    import ceep_amd
    from ceep_amd import adjoint
    from ceep_amd.constants import cfl_timestep_3d
    from ceep_amd.waveforms import modulated_gaussian

    n3, dx3 = 200, 1e-3
    dt3 = cfl_timestep_3d(dx3, dx3, dx3, 0.5) * 0.99
    ring = ceep_amd.phantoms.ring_positions((100, 100), 70, 16)
    pos = [(x, y, 100) for x, y in ring]
    w = modulated_gaussian(T3, dt3, 2e9, 1.5e9)
    freqs = np.linspace(0.5e9, 3e9, 101)

    def solver(seed):
        s = ceep_amd.BatchedFDTD3D(nx=n3, ny=n3, nz=n3, dx=dx3, total_steps=T3,
                                   cpml_thickness=10, source_positions=pos, probe_positions=pos,
                                   frequency=2e9, waveform=w)
        s.set_phantom(ceep_amd.phantoms.DebyeHeadPhantom3D((n3, n3, n3), seed))
        if tuning is not None:
            s.tuning = tuning
        return s

    tgt = solver(1)
    tgt.run()
    S_t = tgt.s_parameters(freqs).s_matrix
    tgt.close()
    s = solver(0)
    if single:
        adjoint.gradient(s, s._eps_r.copy(), s._sigma_e.copy(), freqs, S_t, every=every)
        _device_sync()
        s.close()
        return None
    s._build()
    h = s._handle

    def forward():
        h.reset()
        h.run(copy=False)
        h.dft(freqs)

    forward()
    _device_sync()
    t0 = time.perf_counter()
    forward()
    _device_sync()
    t_fwd = time.perf_counter() - t0
    eps, sig = s._eps_r.copy(), s._sigma_e.copy()
    try:
        adjoint.gradient(s, eps, sig, freqs, S_t, every=every)       # allocates the store
        _device_sync()
        t0 = time.perf_counter()
        adjoint.gradient(s, eps, sig, freqs, S_t, every=every)
        _device_sync()
        t_grad = time.perf_counter() - t0
        slots, extra, _ = h.store_info()
        # the same with the E store of one segment (every segment but the last
        # recomputed in full)
        h.set_tuning("estore_extra_3d", 0)
        adjoint.gradient(s, eps, sig, freqs, S_t, every=every)       # re-sizes the store
        _device_sync()
        t0 = time.perf_counter()
        adjoint.gradient(s, eps, sig, freqs, S_t, every=every)
        _device_sync()
        t_grad1 = time.perf_counter() - t0
        h.set_tuning("estore_extra_3d", -1)
    except Exception as exc:   # e.g. the E store does not fit
        s.close()
        return {"error": f"{type(exc).__name__}: {exc}"}
    box = h.box_info()
    cells = int(box[0] * box[1] * box[2])
    s.close()
    gb = slots * 3 * cells * 16 * 4 / 1e9
    return {"ratio": round(t_grad / t_fwd, 3), "ms_forward": round(t_fwd * 1e3, 1),
            "ms_forward_plus_adjoint": round(t_grad * 1e3, 1), "checkpoint_every": every,
            "e_store_gb": round(gb, 1), "e_store_slots": slots, "e_store_extra_slots": extra,
            "ratio_one_segment_store": round(t_grad1 / t_fwd, 3),
            "note": ("ratio: the E store takes what free HBM allows (e_store_extra_slots "
                     "beyond one segment), so the segment before the last is recomputed "
                     "only in part; ratio_one_segment_store: every segment but the last "
                     "recomputed in full"),
            "workload": ("3D 200^3 Debye+Lorentz head seed 0, 16 TX/RX, 1000 steps; "
                         "L = sum |S - S_target|^2 over 101 freqs 0.5-3 GHz vs head seed 1; "
                         "dL/deps_r and dL/dsigma over the full volume")}


def _roofline(kernel, steps, bytes_launch, launch_ms, summary, dispatches=1,
              launch_ms_one=None, section=None):
    """Roofline object of one kernel: `achieved` / `frac` from the HBM bytes a
    launch moves by rocprofv3 counter (`traffic`, the committed summary) over
    the live mean launch time; the algorithmic figure (every cell's state read
    and written once per step, so blocked launches exceed the peak) beside it.
    K steps over the whole batch may be `dispatches` kernel dispatches (TX
    groups on separate streams, running side by side): the summary's mean
    bytes per dispatch then count that many times, and `launch_ms` is the
    time per K steps of the whole batch.  `launch_ms_one`: the same K steps
    as one dispatch over the whole batch (`one_dispatch`, with that form's
    measured bytes)."""
    per_dispatch = _traffic(kernel, summary, section)
    traffic = per_dispatch * dispatches if per_dispatch else None
    alg = bytes_launch / (launch_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": kernel, "steps_per_launch": steps,
           "dispatches_per_launch": dispatches}
    if launch_ms_one:
        one = _traffic(kernel + " whole batch", summary)
        out["one_dispatch"] = {
            "launch_ms": round(launch_ms_one, 5), "traffic": one,
            "achieved": round(one / (launch_ms_one * 1e-3) / 1e9, 1) if one else None,
            "frac": round(one / (launch_ms_one * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if one else None}
    if traffic:
        ach = traffic / (launch_ms * 1e-3) / 1e9
        out.update({"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic})
    else:
        out.update({"achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                    "traffic": None})
    out.update({"achieved_algorithmic": round(alg, 1),
                "frac_algorithmic": round(alg / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_launch": bytes_launch,
                "launch_ms": round(launch_ms, 5)})
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _traffic(kernel, summary="step2d_debye_600x600x16_traffic.json", section=None):
    """Measured HBM bytes per launch from the committed rocprofv3 counter
    summary, if it was taken for the kernel this run uses (else None);
    `section`: a named part of the summary (a later measurement) instead of
    its top level."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if section is not None:
        d = d.get(section, {})
    # summaries taken before the window kernels named their wavefront count
    # list the four-wavefront kernels without it (same code)
    names = [kernel]
    if kernel.startswith("step2d_mw_kernel<") and kernel.endswith(", 4>"):
        names.append(kernel[:-4] + ">")
    for name in names:
        if "kernels" in d and name in d["kernels"]:
            return d["kernels"][name].get("bytes_per_launch")
        if d.get("kernel") == name:
            return d.get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--single-run", action="store_true",
                    help="one simulation + DFT with the default tuning, no output (for rocprofv3 --pmc)")
    ap.add_argument("--tuning", default=None,
                    help="with --single-run, --single-gradient, --single-run-3d, --single-gradient-3d: "
                         "JSON dict of launch settings applied on top of the defaults")
    ap.add_argument("--single-run-3d", action="store_true",
                    help="one 3D simulation (200^3, Debye + Lorentz, 100 steps), no output (for rocprofv3 --pmc)")
    ap.add_argument("--single-gradient-3d", action="store_true",
                    help="one 3D gradient (200^3, Debye + Lorentz, 100 steps, checkpoints every 20), "
                         "no output (for rocprofv3 --pmc)")
    ap.add_argument("--single-gradient", action="store_true",
                    help="one 2D gradient call on the gradient workload (600^2 disk, 16 TX, 800 steps, "
                         "checkpoints every 50), no output (for rocprofv3 --pmc); --tuning applies")
    ap.add_argument("--single-uneven", action="store_true",
                    help="one uneven batch of the 64 heads with K 8, 64 rows, 4 TX per window, "
                         "eight wavefronts (the tuner picks 4 or 8 TX), no output (for rocprofv3 --pmc)")
    ap.add_argument("--uneven-persistent", action="store_true",
                    help="with --single-uneven: the persistent launch instead of per-round launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy of 256 MiB (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        return
    tun = json.loads(args.tuning) if args.tuning else None
    if args.single_run_3d:
        _bench_3d(steps_timed=1, T3=100, tuning=tun)
        return
    if args.single_gradient_3d:
        _bench_3d_gradient(T3=100, every=20, single=True, tuning=tun)
        return
    if args.single_uneven:
        _bench_uneven(autotune=False, single=True, persistent=args.uneven_persistent)
        return
    if args.single_gradient:
        from ceep_amd import adjoint
        s, eps, sig, freqs, tgt, *_ = _gradient_workload()
        if args.tuning:
            s._handle.apply_tuning(json.loads(args.tuning))
        adjoint.gradient(s, eps, sig, freqs, every=50, target=tgt)
        _device_sync()
        return
    if args.single_run:
        s, *_ = _workload()
        if args.tuning:
            s._handle.apply_tuning(json.loads(args.tuning))
        s._handle.reset()
        s._handle.run(copy=False)
        s._handle.dft(FREQS)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "the batched FDTD path shards TX across GPUs with no exchange and is not "
              "launched on more than one GPU here.", file=sys.stderr)
        sys.exit(2)

    s, ph, pos, w, dt = _workload()
    tuning = None
    if not args.no_autotune:
        from ceep_amd.autotune import autotune_2d, candidates_2d
        # the one-step kernel and the multi-wavefront windows (the one-wavefront
        # windows, never the fastest here, are left to autotune_2d's full list)
        tuning = autotune_2d(s, reps=2, candidates=candidates_2d(one_wavefront=False))
        _progress(f"autotuned 2D launch: {tuning}")
    else:
        from ceep_amd.autotune import tuning_for
        tuning = dict(s.tuning) if s.tuning is not None else tuning_for(
            2, (s.nx, s.ny), s.cpml_n, s.batch, True, False)
    h = s._handle
    box = h.box_info()
    cells = int(box[0] * box[1])
    K = (tuning or {}).get("temporal_depth", 0)
    nw = (tuning or {}).get("forward_waves", 1)
    if K and nw in (4, 8):
        # the head phantom is piecewise constant: material ids unless switched off
        ids = "true" if (tuning or {}).get("material_ids", 1) else "false"
        # the full-depth form (no early exit in the step loop): T is a multiple of K
        full = "true" if T % K == 0 else "false"
        kernel = f"step2d_mw_kernel<{tuning['rows_per_wave'] // nw}, {K}, 1, {ids}, {nw}, {full}>"
        launches = -(-T // K)
        steps_per_launch = K
    elif K:
        kernel = f"step2d_tb_kernel<{tuning['rows_per_wave']}, {K}, 1>"
        launches = -(-T // K)
        steps_per_launch = K
    else:
        kernel = f"step2d_kernel<1, {(tuning or {}).get('batch_per_thread', 4)}>"
        launches = T
        steps_per_launch = 1

    # the blocked kernels split the batch into TX groups of whole chunks of
    # batch_per_wave planes, one group per stream
    dispatches = 1
    if K:
        dispatches = min((tuning or {}).get("tx_streams", 1),
                         -(-B // tuning.get("batch_per_wave", 1)))

    def step():
        h.reset()
        h.run(copy=False)
        return h.dft(FREQS)

    for _ in range(args.warmup):
        step()
    sync = _device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S = step()
    _device_sync()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    value = cells * B * T / (el / args.steps) / 1e9

    # kernel-only timing with HIP events on the library's stream
    ms_run = h.time_runs(3)
    ms_launch = ms_run / launches
    ms_launch_one = None
    if dispatches > 1:
        # the same kernel with the whole batch in one dispatch per K steps:
        # the time a kernel trace of that form reports per dispatch
        h.apply_tuning(dict(tuning, tx_streams=1))
        h.time_runs(1)
        ms_launch_one = h.time_runs(3) / launches
        h.apply_tuning(tuning)
        h.time_runs(1)
    # algorithmic bytes: every time step reads and writes the state once, so
    # a K-step launch is credited K steps' worth (this exceeds the peak when
    # the blocking saves traffic; `traffic` is the measured HBM side)
    bytes_launch = cells * (B * BYTES_PER_CELL_PER_TX + BYTES_PER_CELL_SHARED) * steps_per_launch

    _progress(f"2D workload: {value:.1f} GCell-steps/s")
    adj = None if args.no_adjoint else _adjoint_ratio()
    if adj:
        _progress(f"gradient ratio {adj['ratio']}")
    b3 = None if args.no_3d else _bench_3d()
    if b3:
        _progress(f"3D: {b3['value']} GCell-steps/s")
    g3 = None if (args.no_3d or args.no_adjoint) else _bench_3d_gradient()
    if g3:
        _progress(f"3D gradient: {g3}")
    bu = None if args.no_uneven else _bench_uneven(autotune=not args.no_autotune)
    if bu:
        _progress(f"uneven batch: {bu['value']} GCell-steps/s")
    if not args.no_cpu_baseline:
        _progress("timing the CPU baseline")

    cpu = cpu_omp = None
    if not args.no_cpu_baseline:
        cpu = _cpu_baseline(ph, pos, w, dt, args.cpu_steps)
        cpu_omp = _cpu_openmp(ph, pos, w, dt)

    line = {
        "metric": "GCell-steps/s (600²×16TX, 3D 200³); %HBM peak; fwd+adjoint / fwd time ratio",
        "value": round(value, 3),
        "unit": "GCell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Debye head phantom)",
        "config": {
            "workload": ("2D TMz 600x600 dx=0.5mm Courant 0.5, 15-cell CPML (stepped box "
                         f"{box[0]}x{box[1]}), 16 TX/RX ring r=200, 800 steps, 1-pole Debye head "
                         "seed 0, 301-frequency DFT 0.5-3 GHz to the 16x16 S-matrix"),
            "cells_counted": "stepped box",
            "value_full_grid_convention": round(NX * NY * B * T / (el / args.steps) / 1e9, 3),
            "batch": B, "time_steps": T, "sync": sync, "tuning": tuning,
        },
        "roofline": _roofline(kernel, steps_per_launch, bytes_launch, ms_launch,
                              "step2d_debye_600x600x16_traffic.json", dispatches,
                              ms_launch_one),
        "fwd_adjoint_ratio": adj["ratio"] if adj else None,
        "adjoint": adj,
        "bench_3d": b3,
        "adjoint_3d": g3,
        "bench_uneven": bu,
        "cpu_baseline": cpu,
        "cpu_openmp_reference_point": cpu_omp,
        "s_matrix_shape": list(S.shape),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
