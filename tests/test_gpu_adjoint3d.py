"""Hand-written 3D adjoint (GPU, FP32) with Debye / Lorentz poles vs PyTorch
autograd through the FP64 3D loop and FP64 central differences of the
oracle's loss, on a 26 x 24 x 22 grid (2 TX, 4 RX off the TX cells, a
pulse that peaks inside the run, so the misfit residual is well above FP32
trace accuracy)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import ade, dft, fdtd3d  # noqa: E402
from oracle import waveforms as W  # noqa: E402
from oracle.torch_loop import run_cavity3d_torch, s_matrix_torch  # noqa: E402

pytestmark = pytest.mark.gpu

SH, N, DX, T = (26, 24, 22), 3, 1e-3, 120
SRC = [(10, 11, 10), (15, 12, 12)]
PRB = [(16, 9, 11), (9, 14, 12), (12, 16, 9), (13, 8, 13)]
FREQS = np.linspace(10e9, 45e9, 8)


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _maps(seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(s) for s in SH], indexing="ij")
    c = [s / 2 + rng.uniform(-2, 2) for s in SH]
    m = sum((gi - ci) ** 2 for gi, ci in zip(g, c)) <= rng.uniform(4, 6) ** 2
    eps = np.ones(SH)
    sig = np.zeros(SH)
    eps[m] = rng.uniform(8, 20)
    sig[m] = rng.uniform(0.2, 1.0)
    return eps, sig, m


def _pole_maps(m, kinds):
    out = {}
    if "debye" in kinds:
        out["debye"] = (np.where(m, 30.0, 0.0), np.where(m, 8e-12, 0.0))
    if "lorentz" in kinds:
        out["lorentz"] = (np.where(m, 5.0, 0.0), np.where(m, 2 * np.pi * 20e9, 0.0), 1e9)
    return out


def _oracle_poles(pm, dt):
    poles = []
    if "debye" in pm:
        poles.append(ade.debye_pole(*pm["debye"], dt))
    if "lorentz" in pm:
        poles.append(ade.lorentz_pole(*pm["lorentz"], dt))
    return poles or None


def _solver(amd, eps, sig, pm, w):
    s = amd.BatchedFDTD3D(nx=SH[0], ny=SH[1], nz=SH[2], dx=DX, total_steps=T,
                          cpml_thickness=N, source_positions=SRC, probe_positions=PRB,
                          frequency=30e9, waveform=w)
    if "debye" in pm:
        s.set_debye(eps, pm["debye"][0], pm["debye"][1], sigma_e=sig)
    else:
        s._eps_r[:] = eps
        s._sigma_e[:] = sig
    if "lorentz" in pm:
        s.set_lorentz(*pm["lorentz"])
    return s


@pytest.fixture(scope="module", params=[(), ("debye",), ("lorentz",), ("debye", "lorentz")],
                ids=["plain", "debye", "lorentz", "debye+lorentz"])
def case(request):
    kinds = request.param
    dt = W.dt_3d_reference(DX)
    w = W.modulated_gaussian(T, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps(0)
    eps_t, sig_t, m_t = _maps(1)
    pm = _pole_maps(m, kinds)
    poles = _oracle_poles(pm, dt)
    poles_t = _oracle_poles(_pole_maps(m_t, kinds), dt)
    St = dft.running_dft(fdtd3d.run_cavity(SH, DX, dt, N, SRC, PRB, w, eps_t, sig_t,
                                           poles=poles_t), FREQS, dt)
    e = torch.tensor(eps, requires_grad=True)
    s = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity3d_torch(SH, DX, dt, N, SRC, PRB, w, e, s, poles=poles), FREQS, dt)
    d = S - torch.tensor(St)
    loss = (d.real ** 2 + d.imag ** 2).sum()
    loss.backward()
    cond = float(torch.linalg.norm(d) / torch.linalg.norm(S))
    return dict(kinds=kinds, dt=dt, w=w, eps=eps, sig=sig, pm=pm, poles=poles, St=St,
                loss=float(loss.detach()), ge=e.grad.numpy().copy(), gs=s.grad.numpy().copy(),
                cond=cond)


@pytest.mark.parametrize("every", [1, 17, T])
def test_adjoint3d_matches_autograd(amd, case, every):
    from ceep_amd import adjoint
    c = case
    s = _solver(amd, c["eps"], c["sig"], c["pm"], c["w"])
    assert s.dt == c["dt"]
    loss, ge, gs = adjoint.gradient(s, c["eps"], c["sig"], FREQS, c["St"], every=every)
    ee, es = rel_l2(ge, c["ge"]), rel_l2(gs, c["gs"])
    print(f"3D {c['kinds']} every={every}: residual/|S| {c['cond']:.1e}; loss {loss:.5e} vs "
          f"{c['loss']:.5e}; dL/deps rel-L2 {ee:.2e}, dL/dsigma rel-L2 {es:.2e}", flush=True)
    assert loss == pytest.approx(c["loss"], rel=1e-3)
    assert ee < 1e-3 and es < 1e-3
    # PML cells and the box walls carry no gradient
    assert np.all(ge[:N + 1] == 0) and np.all(ge[:, :, -N:] == 0)


def test_adjoint3d_magnetic_matches_finite_differences(amd):
    """mu_r / sigma_m maps (the magnetic H path) with Debye + Lorentz poles:
    FP64 central differences of the oracle's loss on 4 pixels."""
    from ceep_amd import adjoint
    dt = W.dt_3d_reference(DX)
    w = W.modulated_gaussian(T, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps(0)
    eps_t, sig_t, _ = _maps(1)
    rng = np.random.default_rng(4)
    mu = 1.0 + rng.uniform(0.0, 1.0, SH)
    sm = rng.uniform(0.0, 2e3, SH)
    pm = _pole_maps(m, ("debye", "lorentz"))
    poles = _oracle_poles(pm, dt)
    St = dft.running_dft(fdtd3d.run_cavity(SH, DX, dt, N, SRC, PRB, w, eps_t, sig_t, mu_r=mu,
                                           sigma_m=sm, poles=poles), FREQS, dt)
    s = _solver(amd, eps, sig, pm, w)
    s._mu_r[:] = mu
    s._sigma_m[:] = sm
    _, ge, gs = adjoint.gradient(s, eps, sig, FREQS, St, every=30)

    def loss(ee, ss):
        tr = fdtd3d.run_cavity(SH, DX, dt, N, SRC, PRB, w, ee, ss, mu_r=mu, sigma_m=sm, poles=poles)
        return float(np.sum(np.abs(dft.running_dft(tr, FREQS, dt) - St) ** 2))

    h = 1e-4
    worst = 0.0
    for grad, which in ((ge, "eps"), (gs, "sig")):
        top = np.argsort(np.abs(grad).ravel())[-100:]
        for flat in rng.choice(top, size=2, replace=False):
            a, b, c = np.unravel_index(flat, SH)
            base = eps if which == "eps" else sig
            p, q = base.copy(), base.copy()
            p[a, b, c] += h
            q[a, b, c] -= h
            fd = ((loss(p, sig) - loss(q, sig)) if which == "eps"
                  else (loss(eps, p) - loss(eps, q))) / (2 * h)
            worst = max(worst, abs(grad[a, b, c] - fd) / abs(fd))
    print(f"3D magnetic + Debye + Lorentz: worst relative deviation from FD over 4 pixels {worst:.2e}")
    assert worst < 1e-2
