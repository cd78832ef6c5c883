"""Difference form of the 2D gradient path (``ceep_fdtd_set_reference`` /
``ceep_fdtd_run_residual``): the engine steps E(eps) - E(eps_target) in FP32
with the reference run's source term, so the residual traces -- and the
misfit's dL/dtrace -- are accurate to the residual's own size.

Checked against the FP64 oracle (``oracle.fdtd2d.run_cavity`` of both media,
subtracted in FP64), against a plain run of the target (the reference run's
traces, bit for bit), against FP64 autograd through the torch loop, and on
the edge cases: identical media, a source and probes inside the region where
the media differ, segment lengths that are not multiples of the launch depth."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import dft, fdtd2d  # noqa: E402
from oracle import waveforms as W  # noqa: E402
from oracle.torch_loop import run_cavity_torch, s_matrix_torch  # noqa: E402

pytestmark = pytest.mark.gpu

DX = 0.5e-3


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _solver(amd, n, T, ncp, pos, w, f0):
    s = amd.BatchedFDTD2D(nx=n, ny=n, dx=DX, total_steps=T, cpml_thickness=ncp,
                          source_positions=pos, probe_positions=pos, frequency=f0,
                          courant=0.5, waveform=w)
    s.tuning = {"forward_waves": 4, "temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2}
    return s


def _config1_case(amd):
    """The configuration-1 geometry (200^2, 8 TX = RX on the r = 80 ring, 400
    steps, 2 GHz / 1.5 GHz pulse): disk seed 0 against target disk seed 1."""
    dt = W.dt_2d_courant(DX, 0.5)
    T = 400
    w = W.modulated_gaussian(T, dt, 2e9, 1.5e9)
    pos = amd.phantoms.ring_positions((100, 100), 80, 8)
    eps, sig = amd.phantoms.DiskPhantom2D(200, 200, 0).get_eps_map(2e9)
    eps_t, sig_t = amd.phantoms.DiskPhantom2D(200, 200, 1).get_eps_map(2e9)
    return dict(n=200, T=T, ncp=15, dt=dt, w=w, pos=pos, eps=eps, sig=sig, eps_t=eps_t,
                sig_t=sig_t, f0=2e9)


@pytest.mark.parametrize("every", [50, 47])
def test_residual_traces_match_oracle_difference(amd, every):
    c = _config1_case(amd)
    s = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    s._eps_r[:] = c["eps"]
    s._sigma_e[:] = c["sig"]
    s._build()
    s._handle.set_reference(c["eps_t"], c["sig_t"])
    res, ref = s._handle.run_residual(every)
    args = (c["n"], c["n"], DX, c["dt"], c["ncp"], c["pos"], c["pos"], c["w"])
    t64 = fdtd2d.run_cavity(*args, c["eps"], c["sig"])
    t64_t = fdtd2d.run_cavity(*args, c["eps_t"], c["sig_t"])
    r64 = t64 - t64_t
    er = rel_l2(res, r64)
    et = rel_l2(ref + res, t64)
    eref = rel_l2(ref, t64_t)
    print(f"\nevery={every}: residual |r| / |trace| = {np.linalg.norm(r64) / np.linalg.norm(t64):.1e}; "
          f"residual rel-L2 {er:.2e}; total traces rel-L2 {et:.2e}; reference traces {eref:.2e}")
    assert er <= 1e-4
    assert et <= 1e-5 and eref <= 1e-5
    # the reference run is the plain forward of the target (same kernels, bit for bit)
    p = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    p._eps_r[:] = c["eps_t"]
    p._sigma_e[:] = c["sig_t"]
    p.run()
    assert np.array_equal(ref, p.probe_data)


def test_residual_of_identical_media_is_exactly_zero(amd):
    from ceep_amd import adjoint
    c = _config1_case(amd)
    s = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    freqs = np.linspace(1e9, 2.5e9, 16)
    loss, ge, gs = adjoint.gradient(s, c["eps"], c["sig"], freqs, target=(c["eps"], c["sig"]))
    assert loss == 0.0
    assert not ge.any() and not gs.any()


def test_residual_sources_and_probes_inside_the_difference(amd):
    """A disk of the current medium covers two sources and a probe; T is not a
    multiple of the launch depth; probes on the box wall and outside the box."""
    dt = W.dt_2d_courant(DX, 0.5)
    n, ncp, T = 131, 9, 97
    w = W.modulated_gaussian(T, dt, 6e9, 8e9, delay_factor=3.0)
    pos = [(60, 60), (64, 58), (30, 95), (100, 40)]
    prb = pos + [(62, 61), (ncp, 70), (3, 3)]
    eps = np.ones((n, n)); sig = np.zeros((n, n))
    yy, xx = np.meshgrid(np.arange(n), np.arange(n))
    m = (xx - 62) ** 2 + (yy - 60) ** 2 <= 12 ** 2
    eps[m] = 30.0
    sig[m] = 0.9
    eps_t = np.ones((n, n)); sig_t = np.zeros((n, n))
    mt = (xx - 80) ** 2 + (yy - 75) ** 2 <= 9 ** 2
    eps_t[mt] = 12.0
    sig_t[mt] = 0.4
    s = amd.BatchedFDTD2D(nx=n, ny=n, dx=DX, total_steps=T, cpml_thickness=ncp,
                          source_positions=pos, probe_positions=prb, frequency=6e9,
                          courant=0.5, waveform=w)
    s.tuning = {"forward_waves": 4, "temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 3}
    s._eps_r[:] = eps
    s._sigma_e[:] = sig
    s._build()
    s._handle.set_reference(eps_t, sig_t)
    res, ref = s._handle.run_residual(40)
    args = (n, n, DX, dt, ncp, pos, prb, w)
    r64 = fdtd2d.run_cavity(*args, eps, sig) - fdtd2d.run_cavity(*args, eps_t, sig_t)
    er = rel_l2(res, r64)
    print(f"\nresidual rel-L2 {er:.2e}; probe outside the box: {np.abs(res[:, -1]).max()}")
    assert er <= 1e-4
    assert not res[:, -1].any() and not res[:, -2].any()


def test_residual_gradient_matches_fp64_autograd(amd):
    """dL/d eps_r, dL/d sigma of L = sum |S - S_target|^2 with the target
    given as a medium, against FP64 autograd through the torch loop (96^2
    crop, 4 TX, 240 steps, pulse inside the run)."""
    from ceep_amd import adjoint
    NX, N, T = 96, 8, 240
    freqs = np.linspace(10e9, 25e9, 16)
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 18e9, 20e9, delay_factor=3.0)
    pos = amd.phantoms.ring_positions((48, 48), 30, 4)
    eps, sig = amd.phantoms.DiskPhantom2D(NX, NX, 0, r=(8, 14), max_offset=8).get_eps_map(2e9)
    eps_t, sig_t = amd.phantoms.DiskPhantom2D(NX, NX, 1, r=(8, 14), max_offset=8).get_eps_map(2e9)
    St = dft.running_dft(fdtd2d.run_cavity(NX, NX, DX, dt, N, pos, pos, w, eps_t, sig_t), freqs, dt)
    e = torch.tensor(eps, requires_grad=True)
    sg = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity_torch(NX, NX, DX, dt, N, pos, pos, w, e, sg), freqs, dt)
    d = S - torch.tensor(St)
    L = (d.real ** 2 + d.imag ** 2).sum()
    L.backward()
    s = amd.BatchedFDTD2D(nx=NX, ny=NX, dx=DX, total_steps=T, cpml_thickness=N,
                          source_positions=pos, probe_positions=pos, frequency=18e9,
                          courant=0.5, waveform=w)
    s.tuning = {"forward_waves": 4, "temporal_depth": 6, "rows_per_wave": 32, "batch_per_wave": 2}
    loss, ge, gs = adjoint.gradient(s, eps, sig, freqs, every=50, target=(eps_t, sig_t))
    ee, es = rel_l2(ge, e.grad.numpy()), rel_l2(gs, sg.grad.numpy())
    print(f"\nloss {loss:.6e} vs {float(L):.6e}; dL/deps rel-L2 {ee:.2e}, dL/dsigma rel-L2 {es:.2e}")
    assert loss == pytest.approx(float(L), rel=1e-4)
    assert ee < 1e-3 and es < 1e-3
    # a second call with new materials reuses the reference run
    loss2, ge2, _ = adjoint.gradient(s, eps * 1.01, sig, freqs, every=50, target=(eps_t, sig_t))
    assert loss2 != loss and np.isfinite(ge2).all()


def test_residual_refuses_other_forms(amd):
    c = _config1_case(amd)
    s = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    s.tuning = {"forward_waves": 1, "temporal_depth": 4, "rows_per_wave": 16}
    s._build()
    with pytest.raises(RuntimeError, match="reference medium"):
        s._handle.run_residual(50)
    s._handle.set_reference(c["eps_t"], c["sig_t"])
    with pytest.raises(NotImplementedError, match="multi-wavefront"):
        s._handle.run_residual(50)


@pytest.mark.parametrize("form", ["difference", "s_target"])
def test_fused_gradient_matches_autograd_path(amd, form):
    """``ceep_fdtd_gradient_misfit`` (S-matrix, loss and dL/dtrace on the
    device in FP64 sums) against the autograd path (traces to the host, the
    torch misfit, dL/dtrace back): the same loss and gradients to FP32
    rounding of dL/dtrace, for the difference form and for S_target given as
    data; the traces and S-matrix read afterwards are the same as well."""
    from ceep_amd import adjoint
    c = _config1_case(amd)
    freqs = np.linspace(1e9, 2.5e9, 16)
    s = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    if form == "difference":
        kw = dict(target=(c["eps_t"], c["sig_t"]))
    else:
        args = (c["n"], c["n"], DX, c["dt"], c["ncp"], c["pos"], c["pos"], c["w"])
        St = dft.running_dft(fdtd2d.run_cavity(*args, c["eps_t"], c["sig_t"]), freqs, c["dt"])
        kw = dict(s_target=St)
    L0, ge0, gs0 = adjoint.gradient(s, c["eps"], c["sig"], freqs, every=50, fused=False, **kw)
    tr0 = np.array(s.probe_data)
    S0 = s.s_parameters(freqs).s_matrix
    L1, ge1, gs1 = adjoint.gradient(s, c["eps"], c["sig"], freqs, every=50, **kw)
    tr1 = np.array(s.probe_data)
    S1 = s.s_parameters(freqs).s_matrix
    print(f"\n{form}: loss {L1:.9e} vs {L0:.9e}; dL/deps rel-L2 {rel_l2(ge1, ge0):.2e}, "
          f"dL/dsigma {rel_l2(gs1, gs0):.2e}")
    assert L1 == pytest.approx(L0, rel=1e-9)
    assert rel_l2(ge1, ge0) < 1e-5 and rel_l2(gs1, gs0) < 1e-5
    assert np.array_equal(tr1, tr0)
    assert np.allclose(S1, S0, rtol=0, atol=1e-12 * np.abs(S0).max())
    # zero misfit: zero gradient, no adjoint run
    if form == "difference":
        L2, ge2, gs2 = adjoint.gradient(s, c["eps"], c["sig"], freqs, every=50,
                                        target=(c["eps"], c["sig"]))
        assert L2 == 0.0 and not ge2.any() and not gs2.any()


def test_fused_gradient_rejects_bad_arguments(amd):
    c = _config1_case(amd)
    s = _solver(amd, c["n"], c["T"], c["ncp"], c["pos"], c["w"], c["f0"])
    s._build()
    freqs = np.linspace(1e9, 2.5e9, 4)
    with pytest.raises(ValueError, match="S_target or set a reference"):
        s._handle.gradient_misfit(50, freqs)
    with pytest.raises(ValueError, match="S_target must have shape"):
        s._handle.gradient_misfit(50, freqs, np.zeros((4, 8, 7), dtype=complex))
    # with a reference medium set, an explicit S_target still selects the
    # total-field misfit (the reference is kept)
    s._handle.set_reference(c["eps_t"], c["sig_t"])
    L, ge, _ = s._handle.gradient_misfit(50, freqs, np.zeros((4, 8, 8), dtype=complex))
    assert L > 0 and np.abs(ge).max() > 0


def test_config3_pulse_crop_matches_autograd(amd):
    """Configuration 3's autograd check: a 96 x 96 crop of the configuration-1
    geometry with its pulse (2 GHz, 1.5 GHz bandwidth), 8 TX = RX on a ring,
    400 steps, the misfit at 16 frequencies in 1-2.5 GHz against a target
    disk, checkpoints every 50.  The FP32 difference-form gradient (the fused
    call the bench times) against autograd through the FP64 torch loop."""
    from ceep_amd import adjoint
    NX, N, T = 96, 8, 400
    freqs = np.linspace(1e9, 2.5e9, 16)
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 2e9, 1.5e9)
    pos = amd.phantoms.ring_positions((48, 48), 30, 8)
    eps, sig = amd.phantoms.DiskPhantom2D(NX, NX, 0, r=(8, 14), max_offset=8).get_eps_map(2e9)
    eps_t, sig_t = amd.phantoms.DiskPhantom2D(NX, NX, 1, r=(8, 14), max_offset=8).get_eps_map(2e9)
    St = dft.running_dft(fdtd2d.run_cavity(NX, NX, DX, dt, N, pos, pos, w, eps_t, sig_t), freqs, dt)
    e = torch.tensor(eps, requires_grad=True)
    sg = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity_torch(NX, NX, DX, dt, N, pos, pos, w, e, sg), freqs, dt)
    d = S - torch.tensor(St)
    L = (d.real ** 2 + d.imag ** 2).sum()
    L.backward()
    s = amd.BatchedFDTD2D(nx=NX, ny=NX, dx=DX, total_steps=T, cpml_thickness=N,
                          source_positions=pos, probe_positions=pos, frequency=2e9,
                          courant=0.5, waveform=w)
    loss, ge, gs = adjoint.gradient(s, eps, sig, freqs, every=50, target=(eps_t, sig_t))
    ee, es = rel_l2(ge, e.grad.numpy()), rel_l2(gs, sg.grad.numpy())
    print(f"\nresidual |S - S_t| / |S| = {float(L.detach()) ** 0.5 / float((S.abs() ** 2).sum().detach()) ** 0.5:.1e}; "
          f"loss {loss:.6e} vs {float(L.detach()):.6e}; dL/deps rel-L2 {ee:.2e}, dL/dsigma {es:.2e}")
    assert loss == pytest.approx(float(L.detach()), rel=1e-4)
    assert ee < 1e-3 and es < 1e-3
