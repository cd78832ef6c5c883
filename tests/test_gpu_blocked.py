"""Temporally blocked 2D kernel vs the FP64 oracle and the one-step kernel (GPU).

Odd grid sizes, a step count that is not a multiple of K, a batch that is not
a multiple of the planes per wavefront, sources next to the cavity wall,
probes on the wall, outside the box and duplicated exercise every edge of the
window bookkeeping.
"""
import numpy as np
import pytest

from oracle import ade, fdtd2d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


NX, NY, N, T, DX = 150, 131, 10, 97, 0.5e-3
SRC = [(75, 65), (N + 1, 60), (NX - N - 2, NY - N - 2), (100, 40), (40, 100)]
PRB = [(75, 65), (N + 1, 60), (100, 40), (N, 50), (3, 3), (75, 65), (120, 110), (40, 100)]


def _solver(amd, debye, tuning, T=T):
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 6e9, 8e9, delay_factor=3.0)
    s = amd.BatchedFDTD2D(nx=NX, ny=NY, dx=DX, total_steps=T, cpml_thickness=N,
                          source_positions=SRC, probe_positions=PRB, frequency=6e9,
                          courant=0.5, waveform=w)
    ph = amd.phantoms.DebyeHeadPhantom2D(NX, NY, 3)
    m = ph.get_debye_maps()
    if debye:
        s.set_phantom(ph)
    else:
        s._eps_r[:] = m["eps_inf"] + m["delta_eps"]
        s._sigma_e[:] = m["sigma"]
    s.tuning = tuning
    return s, dt, w, m


def _oracle(s, dt, w, m, debye):
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)] if debye else None
    return fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, s._eps_r, s._sigma_e, poles=poles)


# (K, R, TX planes per window, wavefronts per window); the four-wavefront
# shapes include the default forward (K = 6, R = 32, 2 planes)
CASES = [(1, 16, 1, 1), (2, 24, 2, 1), (3, 32, 1, 1), (4, 16, 1, 1), (4, 24, 3, 1),
         (4, 32, 2, 1), (2, 16, 8, 1),
         (4, 32, 1, 4), (6, 32, 2, 4), (6, 32, 3, 4), (8, 48, 2, 4), (6, 64, 1, 4), (4, 48, 3, 4),
         (6, 64, 2, 8), (4, 32, 3, 8), (6, 48, 1, 8), (8, 64, 2, 8), (4, 64, 5, 8)]


@pytest.mark.parametrize("debye", [False, True])
@pytest.mark.parametrize("K,R,bpw,waves", CASES)
def test_blocked_matches_oracle_and_one_step(amd, debye, K, R, bpw, waves):
    s, dt, w, m = _solver(amd, debye, {"temporal_depth": K, "rows_per_wave": R,
                                       "batch_per_wave": bpw, "forward_waves": waves})
    s.run()
    ref = _oracle(s, dt, w, m, debye)
    err = rel_l2(s.probe_data, ref)
    one, *_ = _solver(amd, debye, {"temporal_depth": 0})
    one.run()
    print(f"K={K} R={R} bpw={bpw} waves={waves} debye={debye}: rel-L2 vs oracle {err:.2e}, "
          f"vs one-step {rel_l2(s.probe_data, one.probe_data):.2e}")
    assert err < TRACE_TOL
    assert rel_l2(s.probe_data, one.probe_data) < 1e-6
    assert np.all(s.probe_data[:, 3:5] == 0)           # wall and CPML probes
    assert np.array_equal(s.probe_data[:, 0], s.probe_data[:, 5])   # duplicate probe


def test_blocked_fields_continuation_and_energy(amd):
    """Odd launch count (T=97, K=4 -> 25 launches) flips the ping-pong parity;
    a second run continues from the fields the first left."""
    tun = {"temporal_depth": 4, "rows_per_wave": 24, "batch_per_wave": 2}
    s, dt, w, m = _solver(amd, True, tun)
    s.run()
    second = s.run()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)]
    ref, st = fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, np.concatenate([w, w]),
                                s._eps_r, s._sigma_e, poles=poles, return_state=True)
    tr = np.array([[second[t][r] for r in second[t]] for t in second])
    assert rel_l2(tr, ref[:, :, T:]) < TRACE_TOL
    sl = (slice(None), slice(N, NX - N + 1), slice(N, NY - N + 1))
    assert rel_l2(s.ez[sl], st["ez"]) < TRACE_TOL
    assert rel_l2(s.hx[sl], st["hx"]) < TRACE_TOL
    assert rel_l2(s.hy[sl], st["hy"]) < TRACE_TOL
    e = (0.5 * W.EPS_0 * np.sum(st["ez"] ** 2) + 0.5 * W.MU_0 * (
        np.sum(st["hx"] ** 2) + np.sum(st["hy"] ** 2))) * DX * DX
    assert s.compute_energy() == pytest.approx(e, rel=1e-5)


@pytest.mark.parametrize("tuning", [
    {"temporal_depth": 3, "rows_per_wave": 16, "batch_per_wave": 1},
    {"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 3, "forward_waves": 4},
])
def test_blocked_run_cpu_wall_sources(amd, tuning):
    nx, ny, T2 = 70, 61, 83
    pos = [(30, 30), (0, 10), (40, 60), (69, 20)]
    prb = [(30, 30), (0, 10), (35, 33), (69, 20)]
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T2, cpml_thickness=5,
                          source_positions=pos, probe_positions=prb, frequency=3e9)
    s.set_material_circle(34, 30, 9, eps_r=7.0, sigma_e=0.3)
    s.tuning = tuning
    out = s.run_cpu()
    ref = fdtd2d.run_cpu(nx, ny, 1e-3, s.dt, pos, prb, W.reference_waveform(T2, s.dt, 3e9),
                         s._eps_r, s._sigma_e)
    tr = np.array([[out[t][r] for r in out[t]] for t in out])
    assert rel_l2(tr, ref) < TRACE_TOL


@pytest.mark.parametrize("case", ["interior", "walls"])
@pytest.mark.parametrize("tuning", [
    {"temporal_depth": 0},
    {"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2, "forward_waves": 4},
])
def test_run_cpu_matches_reference_vectors(amd, golden_dir, case, tuning):
    """run_cpu on the GPU against BatchedFDTD2D.run_cpu() outputs the reference
    wrote (tests/golden/runcpu_*.npz: sources and probes on the walls, in the
    corners, duplicated)."""
    import os
    z = np.load(os.path.join(golden_dir, f"runcpu_{case}.npz"))
    src = [tuple(int(v) for v in p) for p in z["sources"]]
    prb = [tuple(int(v) for v in p) for p in z["probes"]]
    s = amd.BatchedFDTD2D(nx=int(z["nx"]), ny=int(z["ny"]), dx=float(z["dx"]),
                          total_steps=z["traces"].shape[-1], cpml_thickness=4,
                          source_positions=src, probe_positions=prb,
                          frequency=float(z["frequency"]), delay_factor=float(z["delay_factor"]))
    s._eps_r[:] = z["eps_r"]
    s._sigma_e[:] = z["sigma_e"]
    s.tuning = tuning
    out = s.run_cpu()
    tr = np.array([[out[t][r] for r in out[t]] for t in out])
    err = rel_l2(tr, z["traces"])
    print(f"run_cpu {case} {tuning}: rel-L2 vs the reference's vectors {err:.2e}")
    assert s.dt == float(z["dt"])
    assert err < TRACE_TOL


@pytest.fixture(scope="module")
def edge_autograd(amd):
    """FP64 autograd of |S - S_target|^2 through the torch loop on the edge
    geometry (non-dispersive head maps, T = 97)."""
    torch = pytest.importorskip("torch")
    from oracle import dft
    from oracle.torch_loop import run_cavity_torch, s_matrix_torch
    s, dt, w, m = _solver(amd, False, {"temporal_depth": 0})
    freqs = np.linspace(4e9, 9e9, 12)
    eps, sig = s._eps_r.copy(), s._sigma_e.copy()
    eps_t = eps.copy()
    eps_t[60:80, 55:75] *= 1.3
    St = dft.running_dft(fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, eps_t, sig), freqs, dt)
    e = torch.tensor(eps, requires_grad=True)
    sg = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity_torch(NX, NY, DX, dt, N, SRC, PRB, w, e, sg), freqs, dt)
    d = S - torch.tensor(St)
    L = (d.real ** 2 + d.imag ** 2).sum()
    L.backward()
    return dict(freqs=freqs, eps=eps, sig=sig, eps_t=eps_t, St=St,
                ge=e.grad.numpy().copy(), gs=sg.grad.numpy().copy())


@pytest.mark.parametrize("tuning,every", [
    ({"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2, "forward_waves": 4}, 45),
    ({"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 3, "forward_waves": 4,
      "adjoint_batch_per_wave": 3}, 97),
    ({"temporal_depth": 4, "rows_per_wave": 32, "batch_per_wave": 1, "forward_waves": 4,
      "adjoint_temporal_depth": 5, "adjoint_batch_per_wave": 2}, 31),
    ({"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2, "forward_waves": 4,
      "adjoint_temporal_depth": 4, "adjoint_rows_per_wave": 24, "adjoint_batch_per_wave": 8}, 50),
    # eight-wavefront forward and adjoint windows, and the prefetching adjoint
    ({"temporal_depth": 6, "rows_per_wave": 64, "batch_per_wave": 2, "forward_waves": 8,
      "adjoint_waves": 8, "adjoint_rows_per_wave": 48, "adjoint_batch_per_wave": 3}, 50),
    ({"temporal_depth": 4, "rows_per_wave": 32, "batch_per_wave": 3, "forward_waves": 8,
      "adjoint_waves": 8, "adjoint_rows_per_wave": 32, "adjoint_temporal_depth": 5,
      "adjoint_prefetch": 1, "adjoint_batch_per_wave": 2}, 31),
    ({"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2, "forward_waves": 4,
      "adjoint_prefetch": 1, "adjoint_batch_per_wave": 3}, 50),
])
def test_four_wavefront_reversible_adjoint_edges(amd, edge_autograd, tuning, every):
    """The default four-wavefront forward and reversible adjoint on the odd
    grid with wall-adjacent sources, wall / off-box / duplicated probes and
    T = 97 (remainder launches in every segment), against FP64 autograd --
    for both the total-field misfit and the difference form."""
    from ceep_amd import adjoint
    r = edge_autograd
    for mode in ("s_target", "target"):
        s, *_ = _solver(amd, False, tuning)
        if mode == "target":
            s.tuning = {**tuning}
            _, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], r["freqs"], every=every,
                                         target=(r["eps_t"], r["sig"]))
        else:
            _, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], r["freqs"], r["St"], every=every)
        ee, es = rel_l2(ge, r["ge"]), rel_l2(gs, r["gs"])
        print(f"{mode} {tuning} every={every}: dL/deps rel-L2 {ee:.2e}, dL/dsigma {es:.2e}")
        assert ee < 1e-3 and es < 1e-3
        assert not ge[:N + 1].any() and not gs[:, :N + 1].any()


def test_material_ids_bit_identical(amd):
    """Piecewise-constant media read as one byte per cell plus a table in LDS
    (the Debye head has 8 distinct coefficient tuples) give the same bits as
    the full coefficient maps, in a plain run and in an uneven batch."""
    from ceep_amd.uneven import UnevenBatch
    tun = {"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2, "forward_waves": 4}
    out = {}
    for ids in (0, 1):
        s, *_ = _solver(amd, True, {**tun, "material_ids": ids})
        s.run()
        out[ids] = s.probe_data.copy()
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1])
    ub_out = {}
    for ids in (0, 1):
        s1, *_ = _solver(amd, True, {**tun, "material_ids": ids})
        s2, *_ = _solver(amd, True, {**tun, "material_ids": ids}, T=61)
        ub = UnevenBatch([s1, s2], 6, 32, 2, 4)
        ub.run()
        ub_out[ids] = (s1.probe_data.copy(), s2.probe_data.copy())
        ub.close()
    assert np.array_equal(ub_out[0][0], ub_out[1][0]) and np.array_equal(ub_out[0][1], ub_out[1][1])
    assert np.array_equal(ub_out[1][0], out[1])


@pytest.mark.parametrize("streams", [2, 4])
@pytest.mark.parametrize("tuning", [
    {"temporal_depth": 6, "rows_per_wave": 64, "batch_per_wave": 2, "forward_waves": 8},
    {"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 3, "forward_waves": 4},
    {"temporal_depth": 3, "rows_per_wave": 24, "batch_per_wave": 1, "forward_waves": 1},
])
def test_tx_groups_on_streams_bit_identical(amd, streams, tuning):
    """TX groups of the blocked forward on 2 or 4 streams (tuning key 27; the
    batch of 5 TX split into whole chunks, some groups empty or ragged):
    traces and final fields bit-identical to one stream, with and without
    the captured graph and for a second run continuing the first."""
    out = {}
    for ts in (1, streams):
        for graph in (1, 0):
            s, *_ = _solver(amd, True, {**tuning, "tx_streams": ts, "graph": graph})
            s.run()
            first = s.probe_data.copy()
            s.run()
            out[(ts, graph)] = (first, s.probe_data.copy(), s.ez.copy(), s.hy.copy())
    ref = out[(1, 1)]
    assert np.abs(ref[0]).max() > 0
    for key, val in out.items():
        for a, b in zip(val, ref):
            assert np.array_equal(a, b), key


@pytest.mark.parametrize("streams", [2, 4])
def test_tx_groups_gradient_bit_identical(amd, edge_autograd, streams):
    """The difference-form forward, its reference run and the reversible
    adjoint with TX groups on 2 or 4 streams: loss and gradients bit-identical
    to one stream (per-chunk gradient partials are summed in the same order)."""
    from ceep_amd import adjoint
    r = edge_autograd
    res = []
    for ts in (1, streams):
        s, *_ = _solver(amd, False, {"temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2,
                                     "forward_waves": 4, "adjoint_batch_per_wave": 2,
                                     "tx_streams": ts})
        res.append(adjoint.gradient(s, r["eps"], r["sig"], r["freqs"], every=31,
                                    target=(r["eps_t"], r["sig"])))
        res.append(adjoint.gradient(s, r["eps"], r["sig"], r["freqs"], r["St"], every=31))
    for a, b in zip(res[2:], res[:2]):
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
