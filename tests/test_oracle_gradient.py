"""The FP64 torch loop (gradient reference) against the NumPy oracle and FP64
central finite differences -- CPU only."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import dft, fdtd2d  # noqa: E402
from oracle import waveforms as W  # noqa: E402
from oracle.torch_loop import run_cavity_torch, s_matrix_torch  # noqa: E402

NX, NY, N, DX, T = 40, 36, 4, 1e-3, 70
SRC = [(15, 14), (24, 20)]
PRB = [(15, 14), (24, 20), (20, 25)]


def _maps(seed):
    rng = np.random.default_rng(seed)
    eps = np.ones((NX, NY))
    sig = np.zeros((NX, NY))
    x, y = np.meshgrid(np.arange(NX), np.arange(NY), indexing="ij")
    m = (x - 20) ** 2 + (y - 18) ** 2 <= 49
    eps[m] = rng.uniform(5, 20)
    sig[m] = rng.uniform(0.2, 1.0)
    return eps, sig


def _setup():
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    return dt, w


def test_torch_loop_matches_numpy_oracle():
    dt, w = _setup()
    eps, sig = _maps(0)
    a = fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, eps, sig)
    b = run_cavity_torch(NX, NY, DX, dt, N, SRC, PRB, w, torch.tensor(eps),
                         torch.tensor(sig)).numpy()
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    freqs = np.linspace(10e9, 25e9, 5)
    Sa = dft.running_dft(a, freqs, dt)
    Sb = s_matrix_torch(torch.tensor(a), freqs, dt).numpy()
    assert np.allclose(Sa, Sb, rtol=1e-12, atol=1e-30)


def _loss_np(eps, sig, dt, w, freqs, St):
    S = dft.running_dft(fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, eps, sig), freqs, dt)
    return float(np.sum(np.abs(S - St) ** 2))


def test_autograd_matches_central_differences():
    dt, w = _setup()
    eps, sig = _maps(0)
    eps_t, sig_t = _maps(1)
    freqs = np.linspace(10e9, 25e9, 6)
    St = dft.running_dft(fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, eps_t, sig_t), freqs, dt)
    e = torch.tensor(eps, requires_grad=True)
    s = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity_torch(NX, NY, DX, dt, N, SRC, PRB, w, e, s), freqs, dt)
    d = S - torch.tensor(St)
    (d.real ** 2 + d.imag ** 2).sum().backward()
    rng = np.random.default_rng(5)
    cells = [(int(a), int(b)) for a, b in zip(rng.integers(14, 27, 6), rng.integers(12, 25, 6))]
    for (a, b) in cells:
        for grad, arr, h in ((e.grad, eps, 1e-4), (s.grad, sig, 1e-4)):
            p, m = arr.copy(), arr.copy()
            p[a, b] += h
            m[a, b] -= h
            if arr is eps:
                fd = (_loss_np(p, sig, dt, w, freqs, St) - _loss_np(m, sig, dt, w, freqs, St)) / (2 * h)
            else:
                fd = (_loss_np(eps, p, dt, w, freqs, St) - _loss_np(eps, m, dt, w, freqs, St)) / (2 * h)
            g = float(grad[a, b])
            assert abs(g - fd) <= 1e-4 * max(abs(fd), 1e-3 * float(grad.abs().max())), (a, b, g, fd)


def _maps3(shape, seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    m = sum((gi - s / 2) ** 2 for gi, s in zip(g, shape)) <= 25
    eps = np.ones(shape)
    sig = np.zeros(shape)
    eps[m] = rng.uniform(5, 20)
    sig[m] = rng.uniform(0.2, 1.0)
    return eps, sig, m


def _poles3(shape, m, dt):
    from oracle import ade
    return [ade.debye_pole(np.where(m, 30.0, 0.0), np.where(m, 8e-12, 0.0), dt),
            ade.lorentz_pole(np.where(m, 5.0, 0.0), np.where(m, 2 * np.pi * 20e9, 0.0), 1e9, dt)]


SH3, N3, T3, DX3 = (18, 17, 16), 3, 50, 1e-3
SRC3 = [(7, 8, 8), (11, 9, 7)]
PRB3 = [(10, 6, 9), (6, 11, 8)]


def test_torch_loop3d_matches_numpy_oracle():
    from oracle import fdtd3d
    from oracle.torch_loop import run_cavity3d_torch
    dt = W.dt_3d_reference(DX3)
    w = W.modulated_gaussian(T3, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps3(SH3, 0)
    poles = _poles3(SH3, m, dt)
    a = fdtd3d.run_cavity(SH3, DX3, dt, N3, SRC3, PRB3, w, eps, sig, poles=poles)
    b = run_cavity3d_torch(SH3, DX3, dt, N3, SRC3, PRB3, w, torch.tensor(eps), torch.tensor(sig),
                           poles=poles).numpy()
    assert np.abs(a).max() > 0
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_autograd3d_matches_central_differences():
    from oracle import fdtd3d
    from oracle.torch_loop import run_cavity3d_torch
    dt = W.dt_3d_reference(DX3)
    w = W.modulated_gaussian(T3, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps3(SH3, 0)
    eps_t, sig_t, _ = _maps3(SH3, 1)
    poles = _poles3(SH3, m, dt)
    freqs = np.linspace(10e9, 50e9, 5)
    St = dft.running_dft(fdtd3d.run_cavity(SH3, DX3, dt, N3, SRC3, PRB3, w, eps_t, sig_t,
                                           poles=poles), freqs, dt)
    e = torch.tensor(eps, requires_grad=True)
    s = torch.tensor(sig, requires_grad=True)
    S = s_matrix_torch(run_cavity3d_torch(SH3, DX3, dt, N3, SRC3, PRB3, w, e, s, poles=poles),
                       freqs, dt)
    d = S - torch.tensor(St)
    (d.real ** 2 + d.imag ** 2).sum().backward()

    def loss(ee, ss):
        tr = fdtd3d.run_cavity(SH3, DX3, dt, N3, SRC3, PRB3, w, ee, ss, poles=poles)
        return float(np.sum(np.abs(dft.running_dft(tr, freqs, dt) - St) ** 2))

    for (a, b, c) in [(9, 8, 8), (8, 9, 7), (10, 8, 9)]:
        for grad, arr in ((e.grad, eps), (s.grad, sig)):
            h = 1e-4
            p, mm = arr.copy(), arr.copy()
            p[a, b, c] += h
            mm[a, b, c] -= h
            fd = ((loss(p, sig) - loss(mm, sig)) if arr is eps else (loss(eps, p) - loss(eps, mm))) / (2 * h)
            g = float(grad[a, b, c])
            assert abs(g - fd) <= 1e-4 * abs(fd), (a, b, c, g, fd)
