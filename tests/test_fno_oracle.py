"""Pins for the FP64 corrector network reference: central finite differences
of every parameter group and of the input against the hand-derived backward,
plus closed forms of the mode transforms (CPU only)."""
import numpy as np
import pytest

from oracle import fno_oracle as F


def _small(seed=0):
    rng = np.random.default_rng(seed)
    p = F.init_params(rng, C=4, L=3, m=4, Hq=5, scale=1.5)
    for k in ("bl", "b", "bq1", "bq2"):
        p[k] = rng.standard_normal(p[k].shape) * 0.3
    u = rng.standard_normal((2, 3, 8, 12))
    gout = rng.standard_normal((2, 2, 8, 12))
    return p, u, gout


def _loss(p, u, gout):
    out, _ = F.forward(p, u)
    return float(np.sum(out * gout))


@pytest.mark.parametrize("name", ["Wl", "bl", "W", "b", "Q1", "bq1", "Q2", "bq2", "G", "u"])
def test_backward_matches_finite_differences(name):
    p, u, gout = _small()
    _, cache = F.forward(p, u)
    g = F.backward(p, cache, gout)
    rng = np.random.default_rng(7)
    target = u if name == "u" else p[name]
    flat = target.reshape(-1)
    for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
        h = 1e-6
        orig = flat[idx]
        flat[idx] = orig + h
        lp = _loss(p, u, gout)
        flat[idx] = orig - h
        lm = _loss(p, u, gout)
        flat[idx] = orig
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(g[name].reshape(-1)[idx], rel=1e-5, abs=1e-7), (name, idx)


def test_mode_transforms_closed_form():
    H, W, m = 8, 12, 4                      # retained modes k in {-2, -1, 0, 1} per axis
    Fy, Fx = F.dft_mats(H, W, m)
    y, x = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    k = list(F.mode_index(m))
    # cos(2 pi (y/H - 2x/W)) = (e^{i th} + e^{-i th}) / 2; only (ky, kx) = (1, -2) is retained,
    # with weight H*W/2; its partner (-1, 2) lies outside the mode set
    f = np.cos(2 * np.pi * (y / H - 2 * x / W))
    X = F.dft_modes(f, Fy, Fx)
    expect = np.zeros((m, m), complex)
    expect[k.index(1), k.index(-2)] = H * W / 2
    np.testing.assert_allclose(X, expect, atol=1e-9)
    # Re(IDFT) of that single retained coefficient is half the field
    np.testing.assert_allclose(F.idft_modes_real(expect, Fy, Fx), 0.5 * f, atol=1e-12)


def test_gelu_matches_derivative():
    x = np.linspace(-4, 4, 101)
    fd = (F.gelu(x + 1e-6) - F.gelu(x - 1e-6)) / 2e-6
    np.testing.assert_allclose(F.gelu_grad(x), fd, rtol=1e-6, atol=1e-9)


def test_spectral_layer_matches_fft_with_complex_mode_weights():
    """One layer with identity lift and no pointwise term: its pre-activation
    must equal Re(ifft2(R . fft2(u))) restricted to the retained modes, with R
    a complex per-mode weight written as the block [[Re R, -Im R], [Im R, Re R]].
    numpy's FFT fixes the sign, index and normalisation conventions."""
    rng = np.random.default_rng(3)
    C, m, H, W = 3, 4, 8, 12
    p = F.init_params(rng, C=C, L=1, m=m, Hq=4)
    R = rng.standard_normal((m, m, C, C)) + 1j * rng.standard_normal((m, m, C, C))
    p["G"][0] = np.block([[R.real, -R.imag], [R.imag, R.real]])
    p["W"][0] = 0.0
    p["Wl"] = np.eye(C)
    u = rng.standard_normal((2, C, H, W))
    _, cache = F.forward(p, u)
    U = np.fft.fft2(u)
    Y = np.zeros_like(U)
    for a, ky in enumerate(F.mode_index(m)):
        for b, kx in enumerate(F.mode_index(m)):
            Y[:, :, ky % H, kx % W] = np.einsum("oc,bc->bo", R[a, b], U[:, :, ky % H, kx % W])
    np.testing.assert_allclose(cache["pre"][0], np.real(np.fft.ifft2(Y)), atol=1e-10)
