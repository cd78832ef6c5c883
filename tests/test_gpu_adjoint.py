"""Hand-written adjoint (GPU) vs PyTorch autograd through the FP64 loop and
FP64 central finite differences, on a 96 x 96 crop of the configuration-1
geometry (disk phantom, ring of TX/RX)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import dft, fdtd2d  # noqa: E402
from oracle import waveforms as W  # noqa: E402
from oracle.torch_loop import run_cavity_torch, s_matrix_torch  # noqa: E402

pytestmark = pytest.mark.gpu

NX = NY = 96
N, DX, T = 8, 0.5e-3, 240
FREQS = np.linspace(10e9, 25e9, 16)


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def _case(amd):
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 18e9, 20e9, delay_factor=3.0)
    pos = amd.phantoms.ring_positions((48, 48), 30, 4)
    ph = amd.phantoms.DiskPhantom2D(NX, NY, 0, r=(8, 14), max_offset=8)
    tg = amd.phantoms.DiskPhantom2D(NX, NY, 1, r=(8, 14), max_offset=8)
    eps, sig = ph.get_eps_map(2e9)
    eps_t, sig_t = tg.get_eps_map(2e9)
    St = dft.running_dft(fdtd2d.run_cavity(NX, NY, DX, dt, N, pos, pos, w, eps_t, sig_t), FREQS, dt)
    return dt, w, pos, eps, sig, St


def _solver(amd, dt, w, pos):
    return amd.BatchedFDTD2D(nx=NX, ny=NY, dx=DX, total_steps=T, cpml_thickness=N,
                             source_positions=pos, probe_positions=pos, frequency=18e9,
                             courant=0.5, waveform=w)


@pytest.fixture(scope="module")
def reference(amd):
    dt, w, pos, eps, sig, St = _case(amd)
    e = torch.tensor(eps, requires_grad=True)
    s = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity_torch(NX, NY, DX, dt, N, pos, pos, w, e, s), FREQS, dt)
    d = S - torch.tensor(St)
    loss = (d.real ** 2 + d.imag ** 2).sum()
    loss.backward()
    return dict(dt=dt, w=w, pos=pos, eps=eps, sig=sig, St=St, loss=float(loss),
                ge=e.grad.numpy().copy(), gs=s.grad.numpy().copy())


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("every", [50, 1, T])
def test_adjoint_matches_autograd(amd, reference, every):
    from ceep_amd import adjoint
    r = reference
    s = _solver(amd, r["dt"], r["w"], r["pos"])
    loss, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], FREQS, r["St"], every=every)
    assert loss == pytest.approx(r["loss"], rel=1e-4)
    ee, es = rel_l2(ge, r["ge"]), rel_l2(gs, r["gs"])
    print(f"every={every}: loss {loss:.4e} vs {r['loss']:.4e}; dL/deps rel-L2 {ee:.2e}, "
          f"dL/dsigma rel-L2 {es:.2e}")
    assert ee < 1e-2 and es < 1e-2
    assert np.all(ge[:N + 1] == 0) and np.all(gs[:, :N + 1] == 0)   # CPML and wall cells


@pytest.mark.parametrize("every", [50, 7])
def test_adjoint_fp64_forward_matches_autograd(amd, reference, every):
    """The FP64 forward (bitwise the oracle's FP64 traces) with the FP32 adjoint."""
    from ceep_amd import adjoint
    r = reference
    s = _solver(amd, r["dt"], r["w"], r["pos"])
    loss, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], FREQS, r["St"], every=every,
                                    fp64_forward=True)
    ref_tr = fdtd2d.run_cavity(NX, NY, DX, r["dt"], N, r["pos"], r["pos"], r["w"], r["eps"], r["sig"])
    d = float(np.abs(s.probe_data - ref_tr).max() / np.abs(ref_tr).max())
    ee, es = rel_l2(ge, r["ge"]), rel_l2(gs, r["gs"])
    print(f"fp64 forward every={every}: traces max |diff| / max |ref| {d:.1e} "
          f"(bitwise: {np.array_equal(s.probe_data, ref_tr)}); loss {loss:.6e} vs {r['loss']:.6e}; "
          f"dL/deps rel-L2 {ee:.2e}, dL/dsigma rel-L2 {es:.2e}")
    assert d <= 1e-13
    assert loss == pytest.approx(r["loss"], rel=1e-10)
    assert ee < 1e-4 and es < 1e-4


@pytest.mark.parametrize("every", [50, 7, 240])
def test_fp64_forward_blocked_equals_one_step(amd, reference, every):
    """The blocked FP64 forward (4 steps per launch on four-wavefront
    windows) reproduces the one-step FP64 kernel bit for bit, checkpoints and
    segment-length remainders included."""
    r = reference
    out = {}
    for depth in (0, 4):
        s = _solver(amd, r["dt"], r["w"], r["pos"])
        s._eps_r[:] = r["eps"]
        s._sigma_e[:] = r["sig"]
        s.tuning = {"temporal_depth": depth, "fp64_forward": 1}
        s._build()
        out[depth] = s._handle.run_checkpointed(every)
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[4])


@pytest.mark.parametrize("tuning", [
    {"temporal_depth": 0},
    {"temporal_depth": 4, "rows_per_wave": 16, "batch_per_wave": 4},
    {"temporal_depth": 3, "rows_per_wave": 24, "batch_per_wave": 1, "adjoint_temporal_depth": 3},
    {"temporal_depth": 2, "rows_per_wave": 32, "batch_per_wave": 3, "adjoint_temporal_depth": 2},
    {"temporal_depth": 1, "rows_per_wave": 16, "batch_per_wave": 2, "adjoint_temporal_depth": 1},
    {"temporal_depth": 4, "rows_per_wave": 16, "batch_per_wave": 4, "adjoint_reversible": 0},
    {"temporal_depth": 3, "rows_per_wave": 24, "batch_per_wave": 3, "adjoint_reversible": 0,
     "adjoint_temporal_depth": 3},
    {"forward_waves": 4, "temporal_depth": 8, "rows_per_wave": 48, "batch_per_wave": 2},
    {"forward_waves": 4, "temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 3,
     "adjoint_reversible": 0, "adjoint_waves": 4, "adjoint_rows_per_wave": 32},
    {"temporal_depth": 4, "rows_per_wave": 24, "batch_per_wave": 2,
     "adjoint_rows_per_wave": 24, "adjoint_batch_per_wave": 3, "xcd_order": 0},
    {"temporal_depth": 3, "rows_per_wave": 16, "batch_per_wave": 5,
     "adjoint_rows_per_wave": 32, "adjoint_batch_per_wave": 1, "xcd_order": 1},
    {"temporal_depth": 4, "adjoint_waves": 4, "adjoint_rows_per_wave": 32, "adjoint_batch_per_wave": 3},
    {"temporal_depth": 3, "adjoint_waves": 4, "adjoint_rows_per_wave": 24, "adjoint_batch_per_wave": 16,
     "adjoint_temporal_depth": 3},
    {"temporal_depth": 2, "adjoint_waves": 4, "adjoint_rows_per_wave": 16, "adjoint_batch_per_wave": 1,
     "xcd_order": 0},
    {"temporal_depth": 4, "adjoint_waves": 4, "adjoint_temporal_depth": 5, "adjoint_batch_per_wave": 5},
    {"temporal_depth": 4, "adjoint_waves": 4, "adjoint_temporal_depth": 5, "adjoint_rows_per_wave": 24},
    {"temporal_depth": 6, "forward_waves": 4, "rows_per_wave": 32, "batch_per_wave": 2,
     "adjoint_waves": 4, "adjoint_temporal_depth": 6, "adjoint_batch_per_wave": 5},
    {"temporal_depth": 4, "adjoint_waves": 4, "adjoint_temporal_depth": 5},
    {"temporal_depth": 6, "forward_waves": 8, "rows_per_wave": 64, "batch_per_wave": 3,
     "adjoint_waves": 8, "adjoint_rows_per_wave": 32, "adjoint_batch_per_wave": 5},
    {"temporal_depth": 4, "forward_waves": 8, "rows_per_wave": 48, "batch_per_wave": 2,
     "adjoint_waves": 8, "adjoint_rows_per_wave": 48, "adjoint_temporal_depth": 4,
     "adjoint_batch_per_wave": 3},
    {"temporal_depth": 4, "adjoint_waves": 8, "adjoint_rows_per_wave": 32, "adjoint_prefetch": 1,
     "adjoint_temporal_depth": 5, "adjoint_batch_per_wave": 7},
    {"temporal_depth": 4, "adjoint_waves": 4, "adjoint_rows_per_wave": 32, "adjoint_prefetch": 1,
     "adjoint_temporal_depth": 6, "adjoint_batch_per_wave": 3},
])
def test_adjoint_kernel_variants(amd, reference, tuning):
    """One-step and blocked kernels, reversible (default) and recompute + E store
    adjoint forms, segment lengths that are not multiples of K, batch not a
    multiple of the planes per wavefront, adjoint launch shapes other than the
    forward's, both workgroup orders, one- and four-wavefront windows."""
    from ceep_amd import adjoint
    r = reference
    s = _solver(amd, r["dt"], r["w"], r["pos"])
    s.tuning = tuning
    _, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], FREQS, r["St"], every=45)
    ee, es = rel_l2(ge, r["ge"]), rel_l2(gs, r["gs"])
    print(f"{tuning}: dL/deps rel-L2 {ee:.2e}, dL/dsigma rel-L2 {es:.2e}")
    assert ee < 1e-3 and es < 1e-3


def test_adjoint_depth5_needs_multi_wavefront_form(amd, reference):
    from ceep_amd import adjoint
    r = reference
    s = _solver(amd, r["dt"], r["w"], r["pos"])
    s.tuning = {"temporal_depth": 4, "adjoint_waves": 1, "adjoint_temporal_depth": 5}
    with pytest.raises(ValueError, match="multi-wavefront"):
        adjoint.gradient(s, r["eps"], r["sig"], FREQS, r["St"], every=45)


def test_adjoint_matches_finite_differences(amd, reference):
    """FP64 central differences (h = 1e-4) on 32 seeded pixels where the
    gradient is not negligible; agreement within 1%."""
    from ceep_amd import adjoint
    r = reference
    s = _solver(amd, r["dt"], r["w"], r["pos"])
    _, ge, gs = adjoint.gradient(s, r["eps"], r["sig"], FREQS, r["St"], every=50)

    def loss(eps, sig):
        tr = fdtd2d.run_cavity(NX, NY, DX, r["dt"], N, r["pos"], r["pos"], r["w"], eps, sig)
        return float(np.sum(np.abs(dft.running_dft(tr, FREQS, r["dt"]) - r["St"]) ** 2))

    rng = np.random.default_rng(3)

    def pick(g):
        # 16 seeded pixels among the 200 largest |gradient| values
        top = np.argsort(np.abs(g).ravel())[-200:]
        sel = rng.choice(top, size=16, replace=False)
        return [np.unravel_index(i, g.shape) for i in sel]

    picks, picks_s = pick(r["ge"]), pick(r["gs"])
    h = 1e-4
    worst = 0.0
    for (a, b) in picks:
        p, m = r["eps"].copy(), r["eps"].copy()
        p[a, b] += h
        m[a, b] -= h
        fd = (loss(p, r["sig"]) - loss(m, r["sig"])) / (2 * h)
        worst = max(worst, abs(ge[a, b] - fd) / abs(fd))
    for (a, b) in picks_s:
        p, m = r["sig"].copy(), r["sig"].copy()
        p[a, b] += h
        m[a, b] -= h
        fd = (loss(r["eps"], p) - loss(r["eps"], m)) / (2 * h)
        worst = max(worst, abs(gs[a, b] - fd) / abs(fd))
    print(f"worst relative deviation from FD over 32 pixels: {worst:.2e}")
    assert worst < 1e-2
