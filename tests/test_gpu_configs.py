"""Error report per configuration at full size: the HIP engine against the FP64
oracle (its C restatement, bitwise equal to the NumPy oracle) on each
configuration's own grid, steps, batch and seeded phantom.  Every test prints
one line, "config N: ...", with the trace rel-L2 over all (tx, rx, t) and the
S-matrix rel-L2 over all (f, rx, tx); bars: traces <= 1e-5, S <= 1e-4.

  1  2D 200^2, dx 0.5 mm, Courant 0.5, 400 steps, 15-cell CPML, 8 TX = RX on
     the r = 80 ring, 2 GHz / 1.5 GHz modulated Gaussian, seeded lossy disk
     (seeds 0-4)
  2  2D 600^2, 800 steps, 16 TX = RX on the r = 200 ring, Debye head seed 0,
     S at 301 frequencies 0.5-3 GHz
  3  adjoint gradient of L = sum |S - S_target|^2 (16 frequencies 1-2.5 GHz,
     target = disk seed 1, checkpoints every 50 steps) against FP64 central
     differences (h = 1e-4) of the oracle's loss: 32 seeded pixels (16 eps_r,
     16 sigma) on the configuration-1 geometry, and 4 pixels on the 600^2,
     16-TX, 800-step gradient workload the bench times
  4  four problems of the uneven-batch generator (one per grid size 256, 384,
     512, 600; steps U[600, 1200]; CPML 12-15) run together as one uneven batch
  5  3D 200^3, dx 1 mm, 1000 steps, 10-cell CPML, 16 Ez TX = RX on the r = 70
     ring at z = 100, Debye + Lorentz head seed 0, S at 101 frequencies; the
     oracle steps TX 0, 9 and 15 (TX planes are independent; 15 sits in the
     partial last round of workgroups) with all 16 RX
"""
import numpy as np
import pytest

from oracle import ade, c_ref, dft

pytestmark = pytest.mark.gpu

DX2 = 0.5e-3


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _dt2(amd):
    from ceep_amd.constants import cfl_timestep_2d
    return cfl_timestep_2d(DX2, DX2, 0.5)


@pytest.mark.timeout(300)
def test_config1_disk_200(amd):
    from ceep_amd.waveforms import modulated_gaussian
    dt = _dt2(amd)
    T = 400
    pos = amd.phantoms.ring_positions((100, 100), 80, 8)
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    freqs = np.linspace(0.5e9, 3e9, 51)
    worst_t = worst_s = 0.0
    for seed in range(5):
        eps, sig = amd.phantoms.DiskPhantom2D(200, 200, seed).get_eps_map(2e9)
        s = amd.BatchedFDTD2D(nx=200, ny=200, dx=DX2, total_steps=T, cpml_thickness=15,
                              source_positions=pos, probe_positions=pos, frequency=2e9,
                              courant=0.5, waveform=w, cpml_R=1e-6)
        s._eps_r[:] = eps
        s._sigma_e[:] = sig
        s.run()
        S = s.s_parameters(freqs).s_matrix
        ref = c_ref.run_cavity2d(200, 200, DX2, dt, 15, pos, pos, w, eps, sig)
        et = rel_l2(s.probe_data, ref)
        es = rel_l2(S, dft.running_dft(ref, freqs, dt))
        worst_t, worst_s = max(worst_t, et), max(worst_s, es)
        assert et <= 1e-5 and es <= 1e-4, (seed, et, es)
    print(f"\nconfig 1: 200x200 B=8 T=400 disk seeds 0-4: traces rel-L2 <= {worst_t:.2e}, "
          f"S (51 freqs) rel-L2 <= {worst_s:.2e}", flush=True)


@pytest.mark.timeout(300)
def test_config2_debye_head_600(amd):
    from ceep_amd.waveforms import modulated_gaussian
    dt = _dt2(amd)
    T = 800
    pos = amd.phantoms.ring_positions((300, 300), 200, 16)
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    ph = amd.phantoms.DebyeHeadPhantom2D(600, 600, 0)
    s = amd.BatchedFDTD2D(nx=600, ny=600, dx=DX2, total_steps=T, cpml_thickness=15,
                          source_positions=pos, probe_positions=pos, frequency=2e9,
                          courant=0.5, waveform=w, cpml_R=1e-6)
    s.set_phantom(ph)
    s.run()
    freqs = np.linspace(0.5e9, 3e9, 301)
    S = s.s_parameters(freqs).s_matrix
    m = ph.get_debye_maps()
    ref = c_ref.run_cavity2d(600, 600, DX2, dt, 15, pos, pos, w, m["eps_inf"], m["sigma"],
                             poles=[ade.debye_pole(m["delta_eps"], m["tau"], dt)])
    et = rel_l2(s.probe_data, ref)
    es = rel_l2(S, dft.running_dft(ref, freqs, dt))
    print(f"\nconfig 2: 600x600 B=16 T=800 Debye head: traces rel-L2 {et:.2e}, "
          f"S (301 freqs) rel-L2 {es:.2e}", flush=True)
    assert et <= 1e-5 and es <= 1e-4


def _fd_check(amd, nx, T, n, pos, disk_kw, n_pix, seed_pick, mode="target"):
    """Adjoint gradient vs FP64 central differences on n_pix seeded pixels
    (half eps_r, half sigma) drawn from the 400 largest |gradient| cells.

    The misfit residual of these workloads is 5e-6 (200^2) and 3e-10 (600^2)
    of |S| (TX = RX probes see the source pulse itself), below what two FP32
    totals resolve.  mode "target" (the bench's gradient): the target is given
    as a medium and the engine steps the residual itself in FP32; mode
    "fp64": FP64 forward fields, FP32 adjoint, S_target from the oracle."""
    from ceep_amd import adjoint
    from ceep_amd.waveforms import modulated_gaussian
    dt = _dt2(amd)
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    freqs = np.linspace(1e9, 2.5e9, 16)
    eps, sig = amd.phantoms.DiskPhantom2D(nx, nx, 0, **disk_kw).get_eps_map(2e9)
    eps_t, sig_t = amd.phantoms.DiskPhantom2D(nx, nx, 1, **disk_kw).get_eps_map(2e9)
    St = dft.running_dft(c_ref.run_cavity2d(nx, nx, DX2, dt, n, pos, pos, w, eps_t, sig_t),
                         freqs, dt)
    s = amd.BatchedFDTD2D(nx=nx, ny=nx, dx=DX2, total_steps=T, cpml_thickness=n,
                          source_positions=pos, probe_positions=pos, frequency=2e9,
                          courant=0.5, waveform=w, cpml_R=1e-6)
    if mode == "target":
        loss, ge, gs = adjoint.gradient(s, eps, sig, freqs, every=50, target=(eps_t, sig_t))
    else:
        loss, ge, gs = adjoint.gradient(s, eps, sig, freqs, St, every=50, fp64_forward=True)
    print(f"\n  residual |S - S_target| / |S_target| = {np.sqrt(loss) / np.linalg.norm(St):.1e}")

    def fd_loss(e, sg):
        tr = c_ref.run_cavity2d(nx, nx, DX2, dt, n, pos, pos, w, e, sg)
        return float(np.sum(np.abs(dft.running_dft(tr, freqs, dt) - St) ** 2))

    L64 = fd_loss(eps, sig)
    rng = np.random.default_rng(seed_pick)
    h = 1e-4
    worst = 0.0
    for grad, which in ((ge, "eps"), (gs, "sigma")):
        top = np.argsort(np.abs(grad).ravel())[-400:]
        for flat in rng.choice(top, size=n_pix // 2, replace=False):
            a, b = np.unravel_index(flat, grad.shape)
            base = eps if which == "eps" else sig
            p, m = base.copy(), base.copy()
            p[a, b] += h
            m[a, b] -= h
            if which == "eps":
                fd = (fd_loss(p, sig) - fd_loss(m, sig)) / (2 * h)
            else:
                fd = (fd_loss(eps, p) - fd_loss(eps, m)) / (2 * h)
            worst = max(worst, abs(grad[a, b] - fd) / abs(fd))
    return loss, L64, worst


@pytest.mark.timeout(300)
def test_config3_gradient_fd_config1_geometry(amd):
    pos = amd.phantoms.ring_positions((100, 100), 80, 8)
    loss, L64, worst = _fd_check(amd, 200, 400, 15, pos, {}, 32, 5)
    print(f"config 3: 200x200 B=8 T=400, checkpoints every 50, FP32 difference form: "
          f"loss {loss:.4e} (FP64 oracle {L64:.4e}), worst "
          f"relative deviation from FP64 central differences over 32 pixels {worst:.2e}",
          flush=True)
    assert loss == pytest.approx(L64, rel=1e-3)
    assert worst < 1e-2


@pytest.mark.timeout(300)
def test_config3_gradient_fd_config1_geometry_fp64_forward(amd):
    pos = amd.phantoms.ring_positions((100, 100), 80, 8)
    loss, L64, worst = _fd_check(amd, 200, 400, 15, pos, {}, 8, 7, mode="fp64")
    print(f"config 3 (FP64-forward variant): 200x200 B=8 T=400: loss {loss:.4e} "
          f"(FP64 oracle {L64:.4e}), worst relative deviation from FP64 central differences "
          f"over 8 pixels {worst:.2e}", flush=True)
    assert worst < 1e-2


@pytest.mark.timeout(400)
def test_config3_gradient_fd_600(amd):
    """The 600^2, 16-TX, 800-step gradient the bench times (same phantoms,
    target, frequencies, checkpoint interval and difference form)."""
    pos = amd.phantoms.ring_positions((300, 300), 200, 16)
    loss, L64, worst = _fd_check(amd, 600, 800, 15, pos, {"r": (30.0, 90.0), "max_offset": 120.0},
                                 4, 6)
    print(f"config 3: 600x600 B=16 T=800, checkpoints every 50, FP32 difference form: "
          f"loss {loss:.4e} (FP64 oracle {L64:.4e}), worst "
          f"relative deviation from FP64 central differences over 4 pixels {worst:.2e}",
          flush=True)
    assert loss == pytest.approx(L64, rel=1e-3)
    assert worst < 1e-2


@pytest.mark.timeout(300)
def test_config4_uneven_sample(amd):
    from ceep_amd.uneven import UnevenBatch
    from ceep_amd.waveforms import modulated_gaussian
    dt = _dt2(amd)
    rng = np.random.default_rng(11)
    freqs = np.linspace(0.5e9, 3e9, 301)
    solvers, refs = [], []
    for i, g in enumerate((256, 384, 512, 600)):
        steps = int(rng.integers(600, 1201))
        ncp = int(rng.integers(12, 16))
        pos = amd.phantoms.ring_positions((g / 2, g / 2), 200.0 * g / 600.0, 16)
        w = modulated_gaussian(steps, dt, 2e9, 1.5e9)
        ph = amd.phantoms.DebyeHeadPhantom2D(g, g, 2000 + i)
        s = amd.BatchedFDTD2D(nx=g, ny=g, dx=DX2, total_steps=steps, cpml_thickness=ncp,
                              source_positions=pos, probe_positions=pos, frequency=2e9,
                              courant=0.5, waveform=w, cpml_R=1e-6)
        s.set_phantom(ph)
        s._build()
        m = ph.get_debye_maps()
        refs.append(c_ref.run_cavity2d(g, g, DX2, dt, ncp, pos, pos, w, m["eps_inf"], m["sigma"],
                                       poles=[ade.debye_pole(m["delta_eps"], m["tau"], dt)]))
        solvers.append(s)
    ub = UnevenBatch(solvers, 4, 16, 4)
    ub.run()
    worst_t = worst_s = 0.0
    for s, ref in zip(solvers, refs):
        S = s.s_parameters(freqs).s_matrix
        et = rel_l2(s.probe_data, ref)
        es = rel_l2(S, dft.running_dft(ref, freqs, dt))
        worst_t, worst_s = max(worst_t, et), max(worst_s, es)
        assert et <= 1e-5 and es <= 1e-4, (s.nx, s.total_steps, et, es)
    ub.close()
    print(f"\nconfig 4: uneven batch of 256/384/512/600 grids, B=16, steps "
          f"{[s.total_steps for s in solvers]}: traces rel-L2 <= {worst_t:.2e}, "
          f"S (301 freqs) rel-L2 <= {worst_s:.2e}", flush=True)


@pytest.mark.timeout(600)
def test_config5_3d_debye_lorentz_200(amd):
    from ceep_amd.constants import cfl_timestep_3d
    from ceep_amd.waveforms import modulated_gaussian
    n3, dx3, T = 200, 1e-3, 1000
    dt = cfl_timestep_3d(dx3, dx3, dx3, 0.5) * 0.99
    ring = amd.phantoms.ring_positions((100, 100), 70, 16)
    pos = [(x, y, 100) for x, y in ring]
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    ph = amd.phantoms.DebyeHeadPhantom3D((n3, n3, n3), 0)
    s = amd.BatchedFDTD3D(nx=n3, ny=n3, nz=n3, dx=dx3, total_steps=T, cpml_thickness=10,
                          source_positions=pos, probe_positions=pos, frequency=2e9, waveform=w)
    s.set_phantom(ph)
    s.run()
    freqs = np.linspace(0.5e9, 3e9, 101)
    S = s.s_parameters(freqs).s_matrix
    m = ph.get_debye_maps()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt),
             ade.lorentz_pole(m["lorentz_delta_eps"], m["lorentz_omega_0"], m["lorentz_delta"], dt)]
    pick = [0, 9, 15]
    ref = c_ref.run_cavity3d((n3, n3, n3), dx3, dt, 10, [pos[b] for b in pick], pos, w,
                             m["eps_inf"], m["sigma"], poles=poles)
    tr = s.probe_data[pick]
    et = rel_l2(tr, ref)
    es = rel_l2(S[:, :, pick], dft.running_dft(ref, freqs, dt))
    print(f"\nconfig 5: 200^3 B=16 T=1000 Debye+Lorentz head, TX {pick} x 16 RX: traces "
          f"rel-L2 {et:.2e}, S (101 freqs) rel-L2 {es:.2e}", flush=True)
    assert et <= 1e-5 and es <= 1e-4


@pytest.mark.timeout(600)
def test_config5_gradient_fd_2tx(amd, golden_dir):
    """The configuration-5 gradient at full size on a 2-TX sub-batch (TX 0 and
    9 of the ring, all 16 probes; 200^3, 1000 steps, checkpoints every 100,
    Debye + Lorentz head seed 0 against seed 1, 101 frequencies): dL/d eps_r
    and dL/d sigma at four voxels against FP64 central differences (h = 1e-4)
    of the oracle's C restatement, stored with the FP64 target S-matrix and
    loss in tests/golden/fd3d_config5.npz (tests/golden/make_fd3d.py)."""
    import os
    from ceep_amd import adjoint
    from ceep_amd.constants import cfl_timestep_3d
    from ceep_amd.waveforms import modulated_gaussian
    z = np.load(os.path.join(golden_dir, "fd3d_config5.npz"))
    n3, dx3, T = 200, 1e-3, 1000
    dt = cfl_timestep_3d(dx3, dx3, dx3, 0.5) * 0.99
    tx = [tuple(int(v) for v in p) for p in z["tx"]]
    prb = [tuple(int(v) for v in p) for p in z["probes"]]
    w = modulated_gaussian(T, dt, 2e9, 1.5e9)
    s = amd.BatchedFDTD3D(nx=n3, ny=n3, nz=n3, dx=dx3, total_steps=T, cpml_thickness=10,
                          source_positions=tx, probe_positions=prb, frequency=2e9, waveform=w)
    s.set_phantom(amd.phantoms.DebyeHeadPhantom3D((n3, n3, n3), 0))
    St = z["s_target"]
    loss, ge, gs = adjoint.gradient(s, s._eps_r.copy(), s._sigma_e.copy(), z["freqs"], St,
                                    every=100)
    worst = 0.0
    for grad, key in ((ge, "eps"), (gs, "sigma")):
        for flat, fd in zip(z[f"{key}_idx"], z[f"{key}_fd"]):
            v = np.unravel_index(int(flat), (n3, n3, n3))
            dev = abs(grad[v] - fd) / abs(fd)
            worst = max(worst, dev)
            print(f"\n  d L / d {key} at {v}: adjoint {grad[v]:.6e}, FD {fd:.6e}, rel dev {dev:.2e}",
                  end="")
    print(f"\nconfig 5 gradient: 200^3, TX 0 and 9, 1000 steps, checkpoints every 100: loss "
          f"{loss:.4e} (FP64 oracle {float(z['loss']):.4e}), residual |S - S_target| / |S_target| "
          f"= {np.sqrt(float(z['loss'])) / np.linalg.norm(St):.1e}; worst relative deviation "
          f"from FP64 central differences over 4 voxels {worst:.2e}", flush=True)
    assert worst < 1e-2
