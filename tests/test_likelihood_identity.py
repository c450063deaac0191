"""CPU check of the algebra behind the support-length likelihood (numpy only):
-1/2 4 df sum_k |d_k - rfft(h)_k|^2 w_k  ==  -2 df (DD - 2 sum_j h_j g_j + hh),
with g the one-sided inverse transform of d w and hh the M-point sum of
|DFT_M(h)|^2 against the wrapped lag kernel of w (M >= 2 * support, or M = N)."""
import numpy as np
import pytest


def _weights(N, dt, f0=200.0):
    F = N // 2 + 1
    f = np.arange(F) / (N * dt)
    x = f / f0
    return np.where(f > 0, x ** 4 / (1 + x ** 4 + x ** 6), 0.0)


def _support_loglike(h, L, d, w, N, df):
    F = N // 2 + 1
    DD = np.sum(np.abs(d) ** 2 * w)
    Z = d * w
    Z[0] = 2 * Z[0].real
    Z[-1] = 2 * Z[-1].real
    g = 0.5 * np.fft.irfft(Z, n=N) * N          # unnormalised Z2D / 2
    V = np.where((np.arange(F) == 0) | (np.arange(F) == F - 1), w, 0.5 * w).astype(complex)
    c = np.fft.irfft(V, n=N) * N                 # lag kernel
    M = 1
    while M < 2 * L and M < N:
        M *= 2
    M = min(max(M, 2), N)
    cM = np.concatenate([c[:M // 2], c[N - M // 2:]])
    C = np.fft.rfft(cM).real
    W = C * 2 / M
    W[0] /= 2
    W[-1] /= 2
    X = np.fft.rfft(h[:min(L, M)], n=M)
    hh = np.sum(np.abs(X) ** 2 * W)
    return -2 * df * (DD - 2 * np.dot(h[:L], g[:L]) + hh)


@pytest.mark.parametrize("L", [1, 37, 300, 512, 700, 1024])
def test_support_identity(L):
    N, dt = 1024, 1.0 / 2048
    df = 1.0 / (N * dt)
    rng = np.random.default_rng(L)
    w = _weights(N, dt)
    h = np.zeros(N)
    h[:L] = rng.normal(size=L)
    d = rng.normal(size=N // 2 + 1) + 1j * rng.normal(size=N // 2 + 1)
    direct = -0.5 * 4 * df * np.sum(np.abs(d - np.fft.rfft(h)) ** 2 * w)
    assert _support_loglike(h, L, d, w, N, df) == pytest.approx(direct, rel=1e-12)
