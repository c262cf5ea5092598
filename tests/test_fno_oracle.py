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


# ---------------------------------------------------------------- spectral corrector
def _spectral_case(seed=3):
    rng = np.random.default_rng(seed)
    B, Ny, Nx = 3, 12, 14
    window = (2, 3, 7, 9)
    pc = rng.standard_normal((B, Ny, Nx)) + 1j * rng.standard_normal((B, Ny, Nx))
    tgt = np.abs(rng.standard_normal((B, Ny, Nx))) ** 2
    out = 0.2 * rng.standard_normal((B, 2, 7, 9))
    return pc, tgt, out, window


def test_spectral_loss_gradient_matches_finite_differences():
    pc, tgt, out, window = _spectral_case()
    s = 1.7
    _, _, g = F.spectral_loss(out, pc, tgt, window, s)
    rng = np.random.default_rng(11)
    for _ in range(12):
        idx = tuple(int(rng.integers(n)) for n in out.shape)
        h = 1e-4                       # the loss is a quartic in o: central differences are exact to O(h^2)
        op, om = out.copy(), out.copy()
        op[idx] += h
        om[idx] -= h
        num = (F.spectral_loss(op, pc, tgt, window, s)[0] - F.spectral_loss(om, pc, tgt, window, s)[0]) / (2 * h)
        assert g[idx] == pytest.approx(num, rel=1e-6, abs=1e-9)


def test_spectral_loss_special_values():
    pc, tgt, out, window = _spectral_case()
    r0, c0, H, W = window
    win = (slice(None), slice(r0, r0 + H), slice(c0, c0 + W))
    # no correction: the coarse CBED's relative squared error over the window
    loss0, sums0, _ = F.spectral_loss(np.zeros_like(out), pc, tgt, window, 1.0)
    I, T = np.abs(pc[win]) ** 2, tgt[win]
    expect = np.sum((I - T) ** 2, axis=(1, 2)) / np.sum(T ** 2, axis=(1, 2))
    np.testing.assert_allclose(sums0[:, 0] / sums0[:, 1], expect, rtol=1e-12)
    assert loss0 == pytest.approx(expect.mean(), rel=1e-12)
    # the loss is invariant under a common scale s (relative error)
    assert F.spectral_loss(out, pc, tgt, window, 3.0)[0] == pytest.approx(F.spectral_loss(out, pc, tgt, window, 1.0)[0],
                                                                         rel=1e-12)
    # a real gain o = g - 1 with |g|^2 = T / I makes the CBED exact: loss 0
    gain = np.sqrt(T / I) - 1.0
    exact = np.stack([gain, np.zeros_like(gain)], axis=1)
    assert F.spectral_loss(exact, pc, tgt, window, 1.0)[0] < 1e-28


def test_spectral_apply_and_inputs_closed_forms():
    pc, tgt, out, window = _spectral_case()
    r0, c0, H, W = window
    cb0 = F.spectral_apply(np.zeros_like(out), pc, window, scale=2.0)
    np.testing.assert_allclose(cb0, 2.0 * np.abs(pc) ** 2, rtol=1e-15)
    minus = np.stack([-np.ones((3, H, W)), np.zeros((3, H, W))], axis=1)      # o = -1 removes the window
    cb = F.spectral_apply(minus, pc, window)
    assert np.all(cb[:, r0:r0 + H, c0:c0 + W] == 0)
    np.testing.assert_allclose(cb[:, :r0], np.abs(pc[:, :r0]) ** 2, rtol=1e-15)
    # inputs: k^2 on the linspace grid at a hand-picked pixel, thickness channel, vhat channel
    p0 = pc[::-1].copy()
    vh = np.ones((2, 12, 14)) * np.array([3.0, 5.0])[:, None, None]
    u = F.spectral_inputs(pc, p0, window, 8.0, 2.0, 0.5, [40.0, 80.0, 120.0], 0.01, vhat=vh, vidx=[1, 0, 1],
                          vscale=4.0, amplitude=True)
    assert u.shape == (3, 8, H, W)
    ky, kx = 4 * 8.0 / 11 - 4.0, 7 * 8.0 / 13 - 4.0           # pixel (r0 + 2, c0 + 4)
    assert u[1, 4, 2, 4] == pytest.approx(0.5 * (ky ** 2 + kx ** 2), rel=1e-12)
    assert u[2, 5, 0, 0] == pytest.approx(1.2)
    assert u[0, 6, 1, 1] == pytest.approx(np.log1p(20.0)) and u[1, 6, 1, 1] == pytest.approx(np.log1p(12.0))
    assert u[2, 0, 3, 3] == pytest.approx(2.0 * pc[2, r0 + 3, c0 + 3].real)
    assert u[2, 3, 3, 3] == pytest.approx(2.0 * p0[2, r0 + 3, c0 + 3].imag)
    assert u[0, 7, 5, 6] == pytest.approx(2.0 * abs(pc[0, r0 + 5, c0 + 6]))


def test_committed_corrector_has_the_specified_architecture():
    """namsa_amd/data/corrector_c4_f8.npz: 4 Fourier layers, 32 channels, 16 x 16
    modes, 7 inputs (coarse and vacuum spectra, phase, thickness, potential
    spectrum), factor 8 at 256^2; the weight count matches those shapes."""
    import json

    d = np.load("namsa_amd/data/corrector_c4_f8.npz", allow_pickle=False)
    cfg = json.loads(str(d["config"]))
    assert (cfg["L"], cfg["C"], cfg["m"], cfg["Hq"]) == (4, 32, 16, 64)
    assert cfg["factor"] == 8 and cfg["dz_coarse"] == 8.0 and cfg["opts"] == 1
    assert (cfg["Ny"], cfg["Nx"]) == (256, 256) and cfg["window"] == [43, 43, 170, 170]
    C, L, m, Hq, cin = 32, 4, 16, 64, 7
    sizes = [C * cin, C, L * m * m * 4 * C * C, L * C * C, L * C, Hq * C, Hq, 2 * Hq, 2]
    assert d["theta_bf16"].size == sum((n + 63) // 64 * 64 for n in sizes)
    w = (d["theta_bf16"].astype(np.uint32) << 16).view(np.float32)
    assert np.isfinite(w).all() and np.abs(w).max() < 1e3
