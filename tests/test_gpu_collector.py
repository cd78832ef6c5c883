"""``MultistaticDataCollector.record_batched`` on the HIP running DFT: one
batched run fills every transmitter's column, routed by antenna position
(s_params.py:234-353), and matches the oracle's DFT of the oracle's traces."""
import numpy as np
import pytest

from oracle import dft, fdtd2d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu


def test_record_batched_matches_oracle():
    import ceep_amd
    from ceep_amd.sparams import MultistaticDataCollector
    n, T, ncp = 96, 200, 8
    dt = W.dt_2d_courant(0.5e-3, 0.5)
    w = W.modulated_gaussian(T, dt, 6e9, 6e9, delay_factor=3.0)
    ants = ceep_amd.phantoms.ring_positions((48, 48), 30, 6)
    src = [ants[k] for k in (4, 0, 2, 5, 1, 3)]          # TX order differs from the array's
    s = ceep_amd.BatchedFDTD2D(nx=n, ny=n, dx=0.5e-3, total_steps=T, cpml_thickness=ncp,
                               source_positions=src, probe_positions=ants, frequency=6e9,
                               courant=0.5, waveform=w)
    s.set_phantom(ceep_amd.phantoms.DiskPhantom2D(n, n, 3, r=(6, 10), max_offset=6))
    s.run()
    freqs = np.linspace(2e9, 10e9, 21)
    col = MultistaticDataCollector(ants, freqs, dx=0.5e-3)
    col.record_batched(s)
    S = col.get_s_parameters().s_matrix
    ref = fdtd2d.run_cavity(n, n, 0.5e-3, dt, ncp, src, ants, w, s._eps_r, s._sigma_e)
    S_ref = dft.running_dft(ref, freqs, dt)                # (F, RX, TX in source order)
    order = [src.index(a) for a in ants]
    err = float(np.linalg.norm(S - S_ref[:, :, order]) / np.linalg.norm(S_ref))
    print(f"\ncollector from one batched run: S rel-L2 {err:.2e}")
    assert err <= 1e-4
