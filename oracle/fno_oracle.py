"""FP64 reference of the coarse-slice corrector network -- TEST INFRASTRUCTURE ONLY.

The corrector has no counterpart in the reference (it is new in this build):
a Fourier neural operator that maps (coarse exit wave re, im, projected
potential) on an H x W grid to a correction of the exit wave.  This module is
the CPU FP64 definition its GPU kernels are checked against: forward and a
hand-derived backward, pinned by central finite differences
(tests/test_fno_oracle.py).

Network (C channels, L Fourier layers, m x m retained modes, hidden width Hq):
  x0      = Wl u + bl                                   (pointwise, 3 -> C)
  pre_l   = K_l(x_{l-1}) + W_l x_{l-1} + b_l            l = 1..L
  x_l     = gelu(pre_l) for l < L,  x_L = pre_L
  h       = gelu(Q1 x_L + bq1)                          (C -> Hq)
  out     = Q2 h + bq2                                  (Hq -> 2)
  K_l(x)  = Re( IDFT( Y ) ) over the modes k = (ky, kx), ky, kx in
            {-m/2 .. m/2-1}; DFT unnormalised, IDFT with 1/(H W);
            [Re Y; Im Y](k) = G_l(k) [Re X; Im X](k), X = DFT(x), where
            G_l(k) is a real 2C x 2C matrix per mode (a real-linear map of
            the stacked real and imaginary parts; a complex C x C weight R
            is the special case [[Re R, -Im R], [Im R, Re R]]).
  gelu(x) = x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))

The spectral (one-shot) corrector's inputs, loss and application are
restated at the end: spectral_inputs, spectral_loss (pinned by central
finite differences, the zero-correction value and a closed-form gain) and
spectral_apply (pinned by its o = 0 and o = -1 special cases).
"""
from __future__ import annotations

import numpy as np

GELU_C = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * x * x)


def mode_index(m):
    return np.arange(m) - m // 2


def dft_mats(H, W, m):
    """Fy[y, p] = exp(-2 pi i ky(p) y / H), Fx[x, q] = exp(-2 pi i kx(q) x / W)."""
    k = mode_index(m)
    Fy = np.exp(-2j * np.pi * np.outer(np.arange(H), k) / H)
    Fx = np.exp(-2j * np.pi * np.outer(np.arange(W), k) / W)
    return Fy, Fx


def dft_modes(x, Fy, Fx):
    """(..., H, W) real -> (..., m, m) complex: sum_yx x Fy[y,p] Fx[x,q]."""
    return np.einsum("...yx,yp,xq->...pq", x, Fy, Fx, optimize=True)


def idft_modes_real(Y, Fy, Fx):
    """(..., m, m) complex -> (..., H, W) real: Re(sum_pq Y conj(Fy) conj(Fx)) / (H W)."""
    H, W = Fy.shape[0], Fx.shape[0]
    return np.real(np.einsum("...pq,yp,xq->...yx", Y, np.conj(Fy), np.conj(Fx), optimize=True)) / (H * W)


def init_params(rng, C=32, L=4, m=16, Hq=64, scale=1.0):
    s = lambda fan: scale / np.sqrt(fan)
    return {
        "Wl": rng.standard_normal((C, 3)) * s(3), "bl": np.zeros(C),
        "G": rng.standard_normal((L, m, m, 2 * C, 2 * C)) * s(2 * C),
        "W": rng.standard_normal((L, C, C)) * s(C), "b": np.zeros((L, C)),
        "Q1": rng.standard_normal((Hq, C)) * s(C), "bq1": np.zeros(Hq),
        "Q2": rng.standard_normal((2, Hq)) * s(Hq), "bq2": np.zeros(2),
    }


def forward(p, u):
    """u: (B, 3, H, W) real -> out (B, 2, H, W), cache for backward."""
    B, _, H, W = u.shape
    L, m = p["G"].shape[0], p["G"].shape[1]
    C = p["W"].shape[1]
    Fy, Fx = dft_mats(H, W, m)
    cache = {"u": u, "Fy": Fy, "Fx": Fx, "x": [], "X": [], "pre": []}
    x = np.einsum("cj,bjyx->bcyx", p["Wl"], u) + p["bl"][None, :, None, None]
    for l in range(L):
        cache["x"].append(x)
        X = dft_modes(x, Fy, Fx)                                   # (B, C, m, m)
        Xs = np.concatenate([X.real, X.imag], axis=1)              # (B, 2C, m, m)
        Ys = np.einsum("pqoc,bcpq->bopq", p["G"][l], Xs)
        Yh = Ys[:, :C] + 1j * Ys[:, C:]
        pre = idft_modes_real(Yh, Fy, Fx) + np.einsum("oc,bcyx->boyx", p["W"][l], x) \
            + p["b"][l][None, :, None, None]
        cache["X"].append(Xs)
        cache["pre"].append(pre)
        x = gelu(pre) if l < L - 1 else pre
    cache["xL"] = x
    hpre = np.einsum("hc,bcyx->bhyx", p["Q1"], x) + p["bq1"][None, :, None, None]
    h = gelu(hpre)
    out = np.einsum("oh,bhyx->boyx", p["Q2"], h) + p["bq2"][None, :, None, None]
    cache["hpre"], cache["h"] = hpre, h
    return out, cache


def backward(p, cache, gout):
    """Gradients of sum(out * gout) for every parameter and u."""
    Fy, Fx = cache["Fy"], cache["Fx"]
    H, W = Fy.shape[0], Fx.shape[0]
    L = p["G"].shape[0]
    C = p["W"].shape[1]
    g = {}
    g["Q2"] = np.einsum("boyx,bhyx->oh", gout, cache["h"])
    g["bq2"] = gout.sum(axis=(0, 2, 3))
    gh = np.einsum("oh,boyx->bhyx", p["Q2"], gout) * gelu_grad(cache["hpre"])
    g["Q1"] = np.einsum("bhyx,bcyx->hc", gh, cache["xL"])
    g["bq1"] = gh.sum(axis=(0, 2, 3))
    gx = np.einsum("hc,bhyx->bcyx", p["Q1"], gh)
    g["G"] = np.zeros_like(p["G"])
    g["W"] = np.zeros_like(p["W"])
    g["b"] = np.zeros_like(p["b"])
    for l in range(L - 1, -1, -1):
        gpre = gx * gelu_grad(cache["pre"][l]) if l < L - 1 else gx
        x, Xs = cache["x"][l], cache["X"][l]
        g["b"][l] = gpre.sum(axis=(0, 2, 3))
        g["W"][l] = np.einsum("boyx,bcyx->oc", gpre, x)
        # y = Re(IDFT(Yh)): (dL/dRe Yh, dL/dIm Yh) = (Re, Im) of DFT(gpre) / (H W)
        gY = dft_modes(gpre, Fy, Fx) / (H * W)
        gYs = np.concatenate([gY.real, gY.imag], axis=1)
        g["G"][l] = np.einsum("bopq,bcpq->pqoc", gYs, Xs)
        gXs = np.einsum("pqoc,bopq->bcpq", p["G"][l], gYs)
        gX = gXs[:, :C] + 1j * gXs[:, C:]
        # X = DFT(x) for real x: dL/dx = Re(sum gX conj(Fy Fx)) = H W * idft_modes_real(gX)
        gx = idft_modes_real(gX, Fy, Fx) * (H * W) + np.einsum("oc,boyx->bcyx", p["W"][l], gpre)
    g["Wl"] = np.einsum("bcyx,bjyx->cj", gx, cache["u"])
    g["bl"] = gx.sum(axis=(0, 2, 3))
    g["u"] = np.einsum("cj,bcyx->bjyx", p["Wl"], gx)
    return g


# ---------------------------------------------------------------- spectral corrector
# The one-shot corrector acts on the coarse exit spectrum Psi_c (|Psi|^2 is
# the CBED) inside the window [r0, r0+H) x [c0, c0+W) that bounds the
# bandwidth mask, with Psi = Psi_c (1 + o), o = out[0] + i out[1].

def k_linspace(n, kmax):
    """k_j = j kmax / (n - 1) - kmax / 2 (the linspace k-grid of the probe and propagator)."""
    return np.arange(n, dtype=np.float64) * kmax / (n - 1) - kmax / 2.0


def spectral_inputs(psi_c, psi_0, window, kmax, s, phase_scale, thick, thick_scale, vhat=None, vidx=None,
                    vscale=0.0, amplitude=False):
    """Network inputs [B][ch][H][W]: s Psi_c (re, im), s Psi_0 (re, im),
    phase_scale |k|^2 (on sample b's grid of extent kmax[b]), thick_b thick_scale,
    then log1p(vscale |vhat|) and |s Psi_c| when given.  psi_c, psi_0:
    [B][Ny][Nx] complex; vhat: [nv][Ny][Nx]."""
    r0, c0, H, W = window
    B, Ny, Nx = psi_c.shape
    km = np.broadcast_to(np.asarray(kmax, dtype=np.float64), (B,))
    k2 = np.stack([k_linspace(Ny, k)[r0:r0 + H, None] ** 2 + k_linspace(Nx, k)[None, c0:c0 + W] ** 2 for k in km])
    win = (slice(None), slice(r0, r0 + H), slice(c0, c0 + W))
    c = psi_c[win] * s
    z = psi_0[win] * s
    ch = [c.real, c.imag, z.real, z.imag, phase_scale * k2,
          np.broadcast_to((np.asarray(thick, dtype=np.float64) * thick_scale)[:, None, None], c.shape)]
    if vhat is not None:
        idx = np.zeros(B, dtype=int) if vidx is None else np.asarray(vidx)
        ch.append(np.log1p(vscale * np.abs(vhat[idx][win])))
    if amplitude:
        ch.append(np.abs(c))
    return np.stack(ch, axis=1)


def spectral_loss(out, psi_c, target, window, s):
    """mean_b sum (I - F)^2 / sum F^2 with I = |s Psi_c (1 + o)|^2 and F = s^2 target
    over the window; returns (loss, per-sample (num, den), dloss/dout [B][2][H][W])."""
    r0, c0, H, W = window
    win = (slice(None), slice(r0, r0 + H), slice(c0, c0 + W))
    c = psi_c[win]
    o = out[:, 0] + 1j * out[:, 1]
    psi = c * (1.0 + o)
    I = np.abs(psi) ** 2 * s * s
    F = target[win] * s * s
    num = np.sum((I - F) ** 2, axis=(1, 2))
    den = np.sum(F ** 2, axis=(1, 2))
    B = out.shape[0]
    loss = float(np.mean(num / den))
    g = (2.0 * (I - F) / den[:, None, None] / B) * 2.0 * s * s
    z = np.conj(psi) * c                         # d|psi|^2/d(o_re) = 2 Re z, d/d(o_im) = -2 Im z
    grad = np.stack([g * z.real, -g * z.imag], axis=1)
    return loss, np.stack([num, den], axis=1), grad


def spectral_apply(out, psi_c, window, scale=1.0):
    """CBED: |Psi_c (1 + o)|^2 scale in the window, |Psi_c|^2 scale elsewhere."""
    r0, c0, H, W = window
    psi = psi_c.astype(np.complex128).copy()
    o = out[:, 0] + 1j * out[:, 1]
    psi[:, r0:r0 + H, c0:c0 + W] *= (1.0 + o)
    return np.abs(psi) ** 2 * scale
