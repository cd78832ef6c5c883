"""The host-side S-matrix and misfit of the gradient path (ceep_amd.adjoint)
against closed forms: the DFT sum written out in NumPy, and the analytic
dL/dtrace = 2 dt sum_f Re[conj(S - S_t)[f, rx, tx] exp(-i w_f s dt)]."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ceep_amd import adjoint  # noqa: E402


def _direct_s(tr, freqs, dt):
    T = tr.shape[-1]
    e = np.exp(-1j * 2 * np.pi * np.asarray(freqs)[:, None] * np.arange(T)[None, :] * dt)
    return dt * np.einsum("bpt,ft->fpb", tr, e)


@pytest.mark.parametrize("B,P,T,F", [(3, 4, 50, 5), (1, 2, 7, 1), (16, 16, 200, 16)])
def test_s_matrix_and_misfit_match_closed_forms(B, P, T, F):
    rng = np.random.default_rng(B * 100 + T)
    tr = rng.standard_normal((B, P, T))
    dt = 8e-13
    for freqs in (np.linspace(1e9, 2.5e9, F), np.linspace(0.3e9, 4e9, F)):   # cached basis per key
        S = adjoint.s_matrix(torch.from_numpy(tr), freqs, dt).numpy()
        Sd = _direct_s(tr, freqs, dt)
        assert S.shape == (F, P, B)
        assert np.abs(S - Sd).max() <= 1e-12 * np.abs(Sd).max()
        St = Sd * (1 + 0.1 * rng.standard_normal(Sd.shape)) + 1e-3 * np.abs(Sd).max()
        t = torch.from_numpy(tr.copy()).requires_grad_(True)
        L = adjoint.s_misfit(t, freqs, dt, torch.from_numpy(St))
        L.backward()
        D = Sd - St
        assert float(L.detach()) == pytest.approx(float(np.sum(np.abs(D) ** 2)), rel=1e-12)
        e = np.exp(-1j * 2 * np.pi * freqs[:, None] * np.arange(T)[None, :] * dt)   # (F, T)
        g = 2 * dt * np.real(np.einsum("fpb,ft->bpt", np.conj(D), e))
        assert np.linalg.norm(t.grad.numpy() - g) <= 1e-11 * np.linalg.norm(g)


def test_misfit_accepts_a_real_target():
    tr = np.random.default_rng(0).standard_normal((2, 3, 40))
    freqs, dt = np.linspace(1e9, 2e9, 4), 1e-12
    St = np.zeros((4, 3, 2))
    L = adjoint.s_misfit(torch.from_numpy(tr), freqs, dt, torch.from_numpy(St))
    assert float(L) == pytest.approx(float(np.sum(np.abs(_direct_s(tr, freqs, dt)) ** 2)), rel=1e-12)
