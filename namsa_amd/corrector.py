"""Learned coarse-slice corrector on MI355X (no counterpart in the reference).

A thick specimen costs one fused slice step per dz = 1 A of thickness.  The
learned form takes coarse steps of f fine slices (f = 8: one transmission
exp(i sigma V_coarse) and one propagator over 8 A) and corrects each probe's
result once, at the exit: the network maps the coarse exit spectrum Psi_c
(the complex amplitude whose squared modulus is the CBED) to a
multiplicative correction, Psi = Psi_c (1 + o), trained against the fine
path's CBEDs on pairs the plain method generates (SpectralMaterial,
SpectralPairPool, make_spectral_pools; namsa_amd.train_corrector).  The
correction costs one network pass per probe whatever the thickness, so it is
amortised over all the coarse slices it replaces fine ones for.

Network (FnoNet; the FP64 definition and hand-derived backward live in the
oracle and are pinned there by finite differences): lift (cin -> C), L
Fourier layers  x <- gelu(Re IDFT(G(k) [Re; Im] DFT(x)) + W x + b)  over
m x m centred modes with a real 2C x 2C weight per mode, and a 2-layer
pointwise head.  Every product -- the truncated DFTs along y and x, the
per-mode channel mixing, the pointwise layers and all weight gradients -- is
one call of the BF16 MFMA GEMM (namsa_gemm_bf16, FP32 accumulation); GELU
and biases run in its epilogue, the rest in the pointwise kernels of the
same library.

Data layout (HBM, FP32): activations [C][B*H*W] (channel-major, so a
pointwise layer is one [C x C] x [C x n] GEMM and the x-DFT is one
[(C B H) x W] x [W x 2m] GEMM); spectral coefficients [m*m][2C][B] (mode-major,
so the mixing is one batched [2C x 2C] x [2C x B] GEMM over the m*m modes).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._lib import NamsaError, call, ptr, stream_handle


def _torch():
    import torch
    return torch


def _addr(t, off=0):
    return t.data_ptr() + 4 * int(off)


class FnoNet:
    """Parameters, forward and backward of the corrector network on the GPU."""

    def __init__(self, H, W, cin=7, C=32, L=4, m=16, Hq=64, seed=0, device="cuda"):
        torch = _torch()
        if m % 2 or m > min(H, W):
            raise ValueError("m must be even and at most the grid size")
        self.H, self.W, self.cin, self.C, self.L, self.m, self.Hq = H, W, cin, C, L, m, Hq
        self.device = torch.device(device)
        self.shapes = {
            "Wl": (C, cin), "bl": (C,), "G": (L, m, m, 2 * C, 2 * C), "W": (L, C, C), "b": (L, C),
            "Q1": (Hq, C), "bq1": (Hq,), "Q2": (2, Hq), "bq2": (2,),
        }
        self.offsets = {}
        total = 0
        for k, shp in self.shapes.items():
            self.offsets[k] = total
            total += int(np.prod(shp))
            total = (total + 63) // 64 * 64          # 256-byte aligned groups
        self.n_params = total
        self.theta = torch.zeros(total, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.theta)
        self.adam_m = torch.zeros_like(self.theta)
        self.adam_v = torch.zeros_like(self.theta)
        self.adam_step = 0
        self.init_params(seed)
        k = np.arange(m) - m // 2
        ty = 2 * np.pi * np.outer(k, np.arange(H)) / H              # [p][y]
        tx = 2 * np.pi * np.outer(np.arange(W), k) / W              # [x][q]
        fy = np.concatenate([np.cos(ty), np.sin(ty)], 0)             # [(r,p)][y]
        fx = np.concatenate([np.cos(tx), -np.sin(tx)], 1)            # [x][(t,q)]
        self.Fy = torch.from_numpy(fy.astype(np.float32)).to(self.device)
        self.Fx = torch.from_numpy(fx.astype(np.float32)).to(self.device)
        self.B = 0                  # samples the training buffers hold
        self.B_common = 0           # ... the mode-space and input / output buffers
        self.B_fused = 0            # ... the channel-last inference buffers
        self.fx_bf = None
        self._graphs = {}           # B -> captured forward_fused

    # ------------------------------------------------------------------ parameters
    def init_params(self, seed=0):
        rng = np.random.default_rng(seed)
        fan = {"Wl": self.cin, "G": 2 * self.C, "W": self.C, "Q1": self.C, "Q2": self.Hq}
        p = {}
        for k, shp in self.shapes.items():
            p[k] = rng.standard_normal(shp) / np.sqrt(fan[k]) if k in fan else np.zeros(shp)
        # the output layer starts at zero: the network begins as "no correction"
        p["Q2"][:] = 0.0
        self.load(p)

    def view(self, name):
        shp = self.shapes[name]
        o = self.offsets[name]
        return self.theta[o:o + int(np.prod(shp))].view(shp)

    def gview(self, name):
        shp = self.shapes[name]
        o = self.offsets[name]
        return self.grad[o:o + int(np.prod(shp))].view(shp)

    def load(self, params):
        torch = _torch()
        for k, shp in self.shapes.items():
            a = np.asarray(params[k], dtype=np.float32).reshape(shp)
            self.view(k).copy_(torch.from_numpy(a))

    def params_numpy(self):
        return {k: self.view(k).cpu().numpy().astype(np.float64) for k in self.shapes}

    def grads_numpy(self):
        return {k: self.gview(k).cpu().numpy().astype(np.float64) for k in self.shapes}

    def state_dict(self):
        return {"config": dict(H=self.H, W=self.W, cin=self.cin, C=self.C, L=self.L, m=self.m, Hq=self.Hq),
                "theta": self.theta.detach().cpu()}

    def _p(self, name, idx=0):
        o = self.offsets[name]
        if idx:
            o += idx * int(np.prod(self.shapes[name][1:]))
        return _addr(self.theta, o)

    def _g(self, name, idx=0):
        o = self.offsets[name]
        if idx:
            o += idx * int(np.prod(self.shapes[name][1:]))
        return _addr(self.grad, o)

    @property
    def st(self):
        """The current stream (taken at each call, so launches follow torch's stream and graph capture)."""
        return stream_handle()

    # ------------------------------------------------------------------ buffers
    def _ensure_common(self, B):
        """Mode-space buffers and the network's input / output for B samples (both forward paths)."""
        if B <= self.B_common:
            return
        self._graphs.clear()                      # captured graphs hold the addresses about to be replaced
        torch = _torch()
        if self.C * B > 65535:
            raise NamsaError("batch too large for the batched mode transforms")
        H, W, C, L, m = self.H, self.W, self.C, self.L, self.m
        n = B * H * W
        f = dict(dtype=torch.float32, device=self.device)
        self.u = torch.empty((self.cin, n), **f)
        self.out = torch.empty((2, n), **f)
        self.gout = torch.empty((2, n), **f)
        self.Xs = torch.empty((L, m * m, 2 * C, B), **f)
        self.Z = torch.empty((C * B * H, 2 * m), **f)
        self.Pm = torch.empty((C * B, 2 * m, 2 * m), **f)
        self.Ys = torch.empty((m * m, 2 * C, B), **f)
        self.gYs = torch.empty((m * m, 2 * C, B), **f)
        self.E = torch.empty((C * B, 2 * m, 2 * m), **f)
        self.Zp = torch.empty((C * B * H, 2 * m), **f)
        self.sums = torch.empty(2 * B, dtype=torch.float64, device=self.device)
        self.B_common = B

    def ensure(self, B):
        """Buffers of the training forward and backward (FP32 activations) for up to B samples."""
        self._ensure_common(B)
        if B <= self.B:
            return
        torch = _torch()
        H, W, C, L, Hq = self.H, self.W, self.C, self.L, self.Hq
        n = B * H * W
        f = dict(dtype=torch.float32, device=self.device)
        self.xs = [torch.empty((C, n), **f) for _ in range(L + 1)]
        self.pre = [torch.empty((C, n), **f) for _ in range(L - 1)]
        self.hpre = torch.empty((Hq, n), **f)
        self.h = torch.empty((Hq, n), **f)
        self.ga = torch.empty((C, n), **f)
        self.gb = torch.empty((C, n), **f)
        self.gh = torch.empty((Hq, n), **f)
        self.ones = torch.empty(n, **f)
        call("namsa_fill", ptr(self.ones), n, np.float32(1.0), self.st)
        self.B = B

    def fused_ok(self):
        """The channel-last BF16 kernels are compiled for 32 channels, 16 modes per axis, head width 64."""
        return (self.C, self.m, self.Hq) == (32, 16, 64) and self.cin <= 8

    def ensure_fused(self, B):
        """Buffers of the channel-last BF16 inference forward for up to B samples."""
        self._ensure_common(B)
        if B <= self.B_fused:
            return
        self._graphs.clear()
        torch = _torch()
        n = B * self.H * self.W
        self.xa = torch.empty((n, self.C), dtype=torch.bfloat16, device=self.device)
        self.xb = torch.empty((n, self.C), dtype=torch.bfloat16, device=self.device)
        if self.fx_bf is None:
            Wp = (self.W + 31) // 32 * 32
            fx = torch.zeros((Wp, 2 * self.m), dtype=torch.float32, device=self.device)
            fx[:self.W] = self.Fx
            self.Wp = Wp
            self.fx_bf = fx.to(torch.bfloat16).contiguous()
            self.fxT_bf = fx.t().contiguous().to(torch.bfloat16)
        self.B_fused = B

    # ------------------------------------------------------------------ GEMM
    def _gemm(self, A, sA, Bm, sB, Cm, sC, M, N, K, batch=1, split=1, alpha=1.0, beta=0.0, bias=None,
              bias_mode=0, gelu=False, P=None):
        call("namsa_gemm_bf16", A, *sA, Bm, *sB, Cm, *sC, M, N, K, batch, split, np.float32(alpha),
             np.float32(beta), bias, bias_mode, 1 if gelu else 0, P, self.st)

    def _split(self, K):
        return max(1, min(65535, K // 4096))

    def _spectral_forward(self, x, Xs, B, alpha):
        """Xs[pq][(r,c)][b] = alpha * (Re, Im) of the centred m x m DFT modes of x [C][B*H*W]."""
        H, W, C, m = self.H, self.W, self.C, self.m
        m2, CB = 2 * m, C * B
        self._gemm(x, (0, W, 1), _addr(self.Fx), (0, m2, 1), _addr(self.Z), (0, m2, 1),
                   CB * H, m2, W, alpha=alpha)
        self._gemm(_addr(self.Fy), (0, H, 1), _addr(self.Z), (H * m2, m2, 1), _addr(self.Pm),
                   (m2 * m2, m2, 1), m2, m2, H, batch=CB)
        call("namsa_modes_combine", ptr(self.Pm), Xs, C, B, m, self.st)

    def _spectral_inverse(self, Ys, dest, B, alpha, beta, gelu=False, P=None):
        """dest = alpha * Re(sum_modes Y e^{+i k.r}) + beta * dest (GELU optional)."""
        H, W, C, m = self.H, self.W, self.C, self.m
        m2, CB = 2 * m, C * B
        call("namsa_modes_expand", Ys, ptr(self.E), C, B, m, self.st)
        self._gemm(_addr(self.Fy), (0, 1, H), _addr(self.E), (m2 * m2, m2, 1), _addr(self.Zp),
                   (H * m2, m2, 1), H, m2, m2, batch=CB)
        self._gemm(_addr(self.Zp), (0, m2, 1), _addr(self.Fx), (0, 1, m2), dest, (0, W, 1),
                   CB * H, W, m2, alpha=alpha, beta=beta, gelu=gelu, P=P)

    # ------------------------------------------------------------------ passes
    def forward(self, B):
        """out[2][B*H*W] from u[cin][B*H*W] (filled by namsa_corrector_pack)."""
        self.ensure(B)
        H, W, C, L, m, Hq = self.H, self.W, self.C, self.L, self.m, self.Hq
        n = B * H * W
        C2 = 2 * C
        self._gemm(self._p("Wl"), (0, self.cin, 1), ptr(self.u), (0, n, 1), ptr(self.xs[0]), (0, n, 1),
                   C, n, self.cin, bias=self._p("bl"), bias_mode=1)
        for l in range(L):
            x = ptr(self.xs[l])
            Xs = _addr(self.Xs, l * m * m * C2 * B)
            self._spectral_forward(x, Xs, B, 1.0)
            self._gemm(self._p("G", l), (C2 * C2, C2, 1), Xs, (C2 * B, B, 1), ptr(self.Ys), (C2 * B, B, 1),
                       C2, B, C2, batch=m * m)
            self._gemm(self._p("W", l), (0, C, 1), x, (0, n, 1), ptr(self.xs[l + 1]), (0, n, 1), C, n, C,
                       bias=self._p("b", l), bias_mode=1)
            last = l == L - 1
            self._spectral_inverse(ptr(self.Ys), ptr(self.xs[l + 1]), B, 1.0 / (H * W), 1.0,
                                   gelu=not last, P=None if last else ptr(self.pre[l]))
        self._gemm(self._p("Q1"), (0, C, 1), ptr(self.xs[L]), (0, n, 1), ptr(self.h), (0, n, 1), Hq, n, C,
                   bias=self._p("bq1"), bias_mode=1, gelu=True, P=ptr(self.hpre))
        self._gemm(self._p("Q2"), (0, Hq, 1), ptr(self.h), (0, n, 1), ptr(self.out), (0, n, 1), 2, n, Hq,
                   bias=self._p("bq2"), bias_mode=1)
        return self.out

    def forward_fused(self, B):
        """out[2][B*H*W] from u[cin][B*H*W] through the channel-last BF16
        kernels (inference only: nothing is kept for a backward pass).  The
        same network as forward(); activations are rounded to bf16 between
        layers instead of held in FP32."""
        if not self.fused_ok():
            raise NamsaError("the fused forward needs C=32, m=16, Hq=64 and at most 8 inputs")
        self.ensure_fused(B)
        H, W, C, L, m = self.H, self.W, self.C, self.L, self.m
        C2 = 2 * C
        m2, CB = 2 * m, C * B
        st = self.st
        call("namsa_fno_lift", ptr(self.u), self.cin, self._p("Wl"), self._p("bl"), ptr(self.fxT_bf), B, H, W,
             self.Wp, ptr(self.xa), ptr(self.Z), st)
        x_in, x_out = self.xa, self.xb
        for l in range(L):
            Xs = _addr(self.Xs, l * m * m * C2 * B)
            self._gemm(_addr(self.Fy), (0, H, 1), _addr(self.Z), (H * m2, m2, 1), _addr(self.Pm),
                       (m2 * m2, m2, 1), m2, m2, H, batch=CB)
            call("namsa_modes_combine", ptr(self.Pm), Xs, C, B, m, st)
            self._gemm(self._p("G", l), (C2 * C2, C2, 1), Xs, (C2 * B, B, 1), ptr(self.Ys), (C2 * B, B, 1),
                       C2, B, C2, batch=m * m)
            call("namsa_modes_expand", ptr(self.Ys), ptr(self.E), C, B, m, st)
            self._gemm(_addr(self.Fy), (0, 1, H), _addr(self.E), (m2 * m2, m2, 1), _addr(self.Zp),
                       (H * m2, m2, 1), H, m2, m2, batch=CB)
            alpha = np.float32(1.0 / (H * W))
            if l < L - 1:
                call("namsa_fno_layer", ptr(x_in), ptr(self.Zp), ptr(self.fx_bf), ptr(self.fxT_bf),
                     self._p("W", l), self._p("b", l), alpha, B, H, W, self.Wp, ptr(x_out), ptr(self.Z), None, None,
                     None, None, None, st)
                x_in, x_out = x_out, x_in
            else:
                call("namsa_fno_layer", ptr(x_in), ptr(self.Zp), ptr(self.fx_bf), ptr(self.fxT_bf),
                     self._p("W", l), self._p("b", l), alpha, B, H, W, self.Wp, None, None, self._p("Q1"),
                     self._p("bq1"), self._p("Q2"), self._p("bq2"), ptr(self.out), st)
        return self.out

    def forward_graph(self, B):
        """forward_fused(B) replayed from a HIP graph captured on first use for
        this B (about 26 launches per pass: issued one by one from the host
        they cost more time than the kernels take at 256^2).  The graph reads
        u and writes out in place; the weights are read at replay time, so
        loading new parameters needs no recapture."""
        torch = _torch()
        g = self._graphs.get(B)
        if g is None:
            self.forward_fused(B)                 # allocations and one-time tables outside the capture
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.forward_fused(B)
            self._graphs[B] = g
        g.replay()
        return self.out

    def fused_bytes(self, B):
        """Algorithmic HBM bytes of one forward_fused pass on B samples: the
        inputs read (4 cin bytes per pixel), the bf16 activation written by the
        lift, read and written by each of the L - 1 GELU layers and read by the
        last one (64 bytes per pixel each way), the two output channels (8
        bytes), and per image row the 4 KB row-DFT operand each layer's row
        kernels write or read (Z, Zp) plus the mode-space products' read of Z
        and write of Zp."""
        n = B * self.H * self.W
        rows = B * self.H
        act = 2 * self.C
        per_pixel = 4 * self.cin + act + 2 * act * (self.L - 1) + act + 8
        zrow = self.C * 2 * self.m * 4
        return n * per_pixel + rows * zrow * (2 * self.L + 2 * self.L)

    def backward(self, B):
        """grad = d(sum out * gout)/d theta for the last forward of B samples."""
        H, W, C, L, m, Hq = self.H, self.W, self.C, self.L, self.m, self.Hq
        n = B * H * W
        C2 = 2 * C
        sp = self._split(n)
        call("namsa_fill", ptr(self.grad), self.n_params, np.float32(0.0), self.st)
        go, ones = ptr(self.gout), ptr(self.ones)
        self._gemm(go, (0, n, 1), ptr(self.h), (0, 1, n), self._g("Q2"), (0, Hq, 1), 2, Hq, n, split=sp)
        self._gemm(go, (0, n, 1), ones, (0, 1, 0), self._g("bq2"), (0, 1, 1), 2, 1, n, split=sp)
        self._gemm(self._p("Q2"), (0, 1, Hq), go, (0, n, 1), ptr(self.gh), (0, n, 1), Hq, n, 2)
        call("namsa_gelu_backward", ptr(self.gh), ptr(self.hpre), Hq * n, self.st)
        gh = ptr(self.gh)
        self._gemm(gh, (0, n, 1), ptr(self.xs[L]), (0, 1, n), self._g("Q1"), (0, C, 1), Hq, C, n, split=sp)
        self._gemm(gh, (0, n, 1), ones, (0, 1, 0), self._g("bq1"), (0, 1, 1), Hq, 1, n, split=sp)
        ga, gb = self.ga, self.gb
        self._gemm(self._p("Q1"), (0, 1, C), gh, (0, n, 1), ptr(ga), (0, n, 1), C, n, Hq)
        for l in range(L - 1, -1, -1):
            if l < L - 1:
                call("namsa_gelu_backward", ptr(ga), ptr(self.pre[l]), C * n, self.st)
            g = ptr(ga)
            x = ptr(self.xs[l])
            Xs = _addr(self.Xs, l * m * m * C2 * B)
            self._gemm(g, (0, n, 1), ones, (0, 1, 0), self._g("b", l), (0, 1, 1), C, 1, n, split=sp)
            self._gemm(g, (0, n, 1), x, (0, 1, n), self._g("W", l), (0, C, 1), C, C, n, split=sp)
            # (dL/dRe Y, dL/dIm Y) = (Re, Im) DFT(gpre) / (H W)
            self._spectral_forward(g, ptr(self.gYs), B, 1.0 / (H * W))
            self._gemm(ptr(self.gYs), (C2 * B, B, 1), Xs, (C2 * B, 1, B), self._g("G", l), (C2 * C2, C2, 1),
                       C2, C2, B, batch=m * m)
            self._gemm(self._p("G", l), (C2 * C2, 1, C2), ptr(self.gYs), (C2 * B, B, 1), ptr(self.Ys),
                       (C2 * B, B, 1), C2, B, C2, batch=m * m)
            # dL/dx = Re(sum gX e^{+i k.r}) + W^T gpre
            self._spectral_inverse(ptr(self.Ys), ptr(gb), B, 1.0, 0.0)
            self._gemm(self._p("W", l), (0, 1, C), g, (0, n, 1), ptr(gb), (0, n, 1), C, n, C, beta=1.0)
            ga, gb = gb, ga
        g = ptr(ga)
        self._gemm(g, (0, n, 1), ptr(self.u), (0, 1, n), self._g("Wl"), (0, self.cin, 1), C, self.cin, n,
                   split=sp)
        self._gemm(g, (0, n, 1), ones, (0, 1, 0), self._g("bl"), (0, 1, 1), C, 1, n, split=sp)

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.adam_step += 1
        call("namsa_adam", ptr(self.theta), ptr(self.grad), ptr(self.adam_m), ptr(self.adam_v), self.n_params,
             np.float32(lr), np.float32(b1), np.float32(b2), np.float32(eps), self.adam_step, np.float32(wd),
             self.st)

    def flops(self, B):
        """Multiply-add flops (x2) of one forward pass on B samples."""
        H, W, C, L, m, Hq = self.H, self.W, self.C, self.L, self.m, self.Hq
        n = B * H * W
        m2 = 2 * m
        spec = 2 * (C * B * H * W * m2 + C * B * m2 * m2 * H)            # y/x DFT
        spec += 2 * (m * m * 2 * C * 2 * C * B)                            # mixing
        spec += 2 * (C * B * H * m2 * m2 + C * B * H * m2 * W)             # inverse
        return 2 * n * C * self.cin + L * (spec + 2 * n * C * C) + 2 * n * (Hq * C + 2 * Hq)


# ====================================================================== spectral corrector
# One correction per probe, applied to the coarse path's exit spectrum (the
# complex amplitude whose squared modulus is the CBED):
#
#     Psi = Psi_c (1 + o),   o = out[0] + i out[1] = FNO(u)
#
# over the window that bounds the bandwidth mask.  Inputs u (per k-pixel):
# s Psi_c, the probe's exit spectrum s Psi_0 through vacuum of the same
# thickness (the reference the network compares against), the phase
# pi lambda dz_c |k|^2 of one coarse propagator step, the thickness, and the
# magnitude of the projected potential's spectrum (where the specimen
# scatters).  s = sqrt(Ny Nx) puts |s Psi|^2 at mean 1 per pixel (the CBED
# sums to 1).  Training minimises the mean over samples of the CBED's
# relative squared error, sum (I - F)^2 / sum F^2, against the fine path.

SPECTRAL_PHASE_DIV = 10.0      # u[4] = pi lambda dz_c |k|^2 / 10
SPECTRAL_THICK_SCALE = 1.0 / 400.0
SPECTRAL_VSCALE = 100.0        # u[6] = log1p(100 |FFT(sigma V_T)| / (Ny Nx))
OPT_POTENTIAL, OPT_AMPLITUDE = 1, 2


def band_window(mask):
    """(r0, c0, H, W): bounding box of the nonzero bandwidth mask [Ny][Nx] (host array)."""
    rows = np.nonzero(np.asarray(mask).any(axis=1))[0]
    cols = np.nonzero(np.asarray(mask).any(axis=0))[0]
    return int(rows[0]), int(cols[0]), int(rows[-1] - rows[0] + 1), int(cols[-1] - cols[0] + 1)


class SpectralMaterial:
    """One structure prepared for fine (dz) and coarse (f * dz, f in factors)
    stepping and for the spectral corrector's inputs.

    Holds the fine slices t_f, coarse slices t_c[f] (the same site table
    binned f * dz thick), the propagators, the bandwidth window and the fine
    site table (for projected potentials of any thickness).
    """

    def __init__(self, sim, dz=1.0, factors=(8,), bandwidth=1 / 3, probe_dict=None):
        torch = _torch()
        self.sim = sim
        self.dz, self.bandwidth = float(dz), float(bandwidth)
        self.factors = tuple(int(f) for f in factors)
        sim._check_grid()
        if not hasattr(sim, "tables_device"):
            sim.calc_atomic_potentials()
        Ny, Nx = sim._grid
        self.Ny, self.Nx = Ny, Nx
        dev = sim.device
        sim.build_potential_slices(dz)
        self.t_f = sim.potential_slices_device
        self.S_f = int(sim.num_slices)
        sites, offs = sim.site_table(dz)
        self.sites_d = torch.from_numpy(sites.reshape(-1)).to(dev) if len(sites) else None
        self.offsets = offs
        self.pm_f = sim.build_propagator(bandwidth).clone()
        self.t_c, self.pm_c, self.S_c = {}, {}, {}
        for f in self.factors:
            sim.build_potential_slices(dz * f)
            self.t_c[f] = sim.potential_slices_device
            self.S_c[f] = int(sim.num_slices)
            if self.S_c[f] * f > self.S_f:
                raise NamsaError("coarse slices extend past the fine ones")
            self.pm_c[f] = sim.build_propagator(bandwidth).clone()
        self.mask = sim.mask_device.clone()
        self.band = sim.band
        self.window = band_window(self.mask.cpu().numpy())
        if not hasattr(sim, "psi_k_device") or probe_dict is not None:
            sim.build_probe(probe_dict) if probe_dict is not None else sim.build_probe()
        self.norm = float(sim.normalization)
        self.lam = float(sim.Lambda)
        self.ones = torch.ones((1, Ny, Nx), dtype=torch.complex64, device=dev)

    def _slices(self, factor):
        return (self.t_f, self.pm_f) if factor in (None, 1) else (self.t_c[factor], self.pm_c[factor])

    def exit_spectra(self, pos, S, factor, out):
        """Centred exit spectra [B][Ny][Nx] after S slices of dz * factor
        (factor None or 1: the fine slices): exit waves scaled
        1/sqrt(Ny Nx norm), then the forward transform, so that |out|^2 is
        the CBED."""
        B = pos.shape[0]
        t, pm = self._slices(factor)
        pm_arg, flags = self._cbed_pm(factor, pm)
        scale = np.float32(1.0 / math.sqrt(self.Ny * self.Nx * self.norm))
        call("namsa_multislice", ptr(self.sim.psi_k_device), ptr(pos), B, self.Ny, self.Nx,
             np.float32(self.sim.kmax), ptr(t), int(S), pm_arg, self.band[0], self.band[1], B,
             _lib.OUT_EXIT | flags, scale, ptr(out), None, None, 0, None, None, stream_handle())
        call("namsa_fft2d", ptr(out), B, self.Ny, self.Nx, 0, stream_handle())
        return out

    def cbed(self, pos, S, factor, out, group=None):
        """Plain-path CBEDs [B][Ny][Nx] after S slices of dz * factor (the
        fused driver, `group` probes through the slices together; default all)."""
        torch = _torch()
        B = pos.shape[0]
        G = min(B, int(group)) if group else B
        t, pm = self._slices(factor)
        pm_arg, flags = self._cbed_pm(factor, pm)
        ws = torch.empty((G, self.Ny, self.Nx), dtype=torch.complex64, device=out.device)
        call("namsa_multislice", ptr(self.sim.psi_k_device), ptr(pos), B, self.Ny, self.Nx,
             np.float32(self.sim.kmax), ptr(t), int(S), pm_arg, self.band[0], self.band[1], G,
             _lib.OUT_CBED | flags, np.float32(1.0 / self.norm), ptr(ws), ptr(out), None, 0, None, None,
             stream_handle())
        return out

    def _cbed_pm(self, factor, pm):
        """The propagator argument and mode flags of a CBED or exit-wave run:
        where the slice loop tiles the wave, a tiled copy (made once per
        factor)."""
        if not _lib.load().namsa_wave_tiled(self.Ny, self.Nx):
            return ptr(pm), 0
        key = 1 if factor in (None, 1) else factor
        cache = self.__dict__.setdefault("_pm_tiled", {})
        if key not in cache:
            cache[key] = _torch().empty_like(pm)
            call("namsa_tile_plane", ptr(pm), ptr(cache[key]), self.Ny, self.Nx, stream_handle())
        return ptr(cache[key]), _lib.PM_TILED

    def vacuum_spectra(self, pos, thickness, out):
        """Exit spectra of the same probes through vacuum of `thickness`: one
        step with t = 1 and the propagator over the whole thickness."""
        torch = _torch()
        pm_t = torch.empty((self.Ny, self.Nx), dtype=torch.complex64, device=out.device)
        call("namsa_propagator", self.Ny, self.Nx, np.float32(self.sim.kmax),
             np.float32(self.bandwidth * self.sim.kmax), np.float32(self.lam), np.float32(thickness), ptr(pm_t),
             None, stream_handle())
        B = pos.shape[0]
        scale = np.float32(1.0 / math.sqrt(self.Ny * self.Nx * self.norm))
        call("namsa_multislice", ptr(self.sim.psi_k_device), ptr(pos), B, self.Ny, self.Nx,
             np.float32(self.sim.kmax), ptr(self.ones), 1, ptr(pm_t), self.band[0], self.band[1], B,
             _lib.OUT_EXIT, scale, ptr(out), None, None, 0, None, None, stream_handle())
        call("namsa_fft2d", ptr(out), B, self.Ny, self.Nx, 0, stream_handle())
        return out

    def potential_spectrum(self, S, out):
        """Centred spectrum of sigma V over the first S fine slices, [Ny][Nx] complex."""
        torch = _torch()
        sim = self.sim
        Py, Px = sim.table_shape
        v = torch.empty((1, self.Ny, self.Nx), dtype=torch.float32, device=out.device)
        offs = torch.tensor([0, int(self.offsets[int(S)])], dtype=torch.int32, device=out.device)
        call("namsa_build_potential", ptr(sim.tables_device), len(sim.unique_Z), Py, Px, ptr(self.sites_d),
             ptr(offs), 1, self.Ny, self.Nx, ptr(v), stream_handle())
        call("namsa_real_to_complex", ptr(v), 1, self.Ny, self.Nx, np.float32(sim.sigma), 1, ptr(out),
             stream_handle())
        call("namsa_fft2d", ptr(out), 1, self.Ny, self.Nx, 0, stream_handle())
        return out


class SpectralCorrector:
    """The spectral corrector network (an FnoNet over the band window) and its
    input packing, loss and application."""

    def __init__(self, Ny, Nx, window, kmax, lam, dz_coarse, opts=OPT_POTENTIAL, C=32, L=4, m=16, Hq=64,
                 seed=0, device="cuda", factor=8):
        self.Ny, self.Nx = int(Ny), int(Nx)
        self.factor = int(factor)
        self.window = tuple(int(v) for v in window)
        self.kmax, self.lam, self.dz_coarse, self.opts = float(kmax), float(lam), float(dz_coarse), int(opts)
        cin = 6 + (1 if opts & OPT_POTENTIAL else 0) + (1 if opts & OPT_AMPLITUDE else 0)
        _, _, H, W = self.window
        self.net = FnoNet(H, W, cin=cin, C=C, L=L, m=m, Hq=Hq, seed=seed, device=device)
        self.s = float(math.sqrt(self.Ny * self.Nx))
        self.phase_scale = float(math.pi * self.lam * self.dz_coarse / SPECTRAL_PHASE_DIV)
        self.fused = True           # inference through the channel-last BF16 kernels where they apply

    def config(self):
        n = self.net
        return dict(Ny=self.Ny, Nx=self.Nx, window=list(self.window), kmax=self.kmax, lam=self.lam,
                    dz_coarse=self.dz_coarse, factor=self.factor, opts=self.opts, C=n.C, L=n.L, m=n.m, Hq=n.Hq)

    def save(self, path):
        """Weights as bfloat16 bit patterns (the network's operand precision) plus the configuration."""
        import json
        theta = self.net.theta.detach().cpu().numpy().astype(np.float32)
        bits = theta.view(np.uint32)
        rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)   # round to nearest even
        np.savez_compressed(path, theta_bf16=rounded, config=np.array(json.dumps(self.config())))

    @classmethod
    def load(cls, path, device="cuda"):
        import json
        torch = _torch()
        d = np.load(path, allow_pickle=False)
        cfg = json.loads(str(d["config"]))
        obj = cls(cfg["Ny"], cfg["Nx"], cfg["window"], cfg["kmax"], cfg["lam"], cfg["dz_coarse"], cfg["opts"],
                  C=cfg["C"], L=cfg["L"], m=cfg["m"], Hq=cfg["Hq"], device=device, factor=cfg["factor"])
        theta = (d["theta_bf16"].astype(np.uint32) << 16).view(np.float32)
        if theta.size != obj.net.n_params:
            raise NamsaError(f"{path}: {theta.size} weights, the network has {obj.net.n_params}")
        obj.net.theta.copy_(torch.from_numpy(theta))
        return obj

    def pack(self, psi_c, psi_0, vhat, vidx, kmax, thick, B):
        r0, c0, H, W = self.window
        self.net._ensure_common(B)
        nv = int(vhat.shape[0]) if vhat is not None and vhat.dim() == 3 else 1
        call("namsa_spectral_pack", ptr(psi_c), ptr(psi_0), ptr(vhat), ptr(vidx), nv, B, self.Ny, self.Nx, r0, c0, H,
             W, ptr(kmax), np.float32(self.s), np.float32(self.phase_scale), ptr(thick),
             np.float32(SPECTRAL_THICK_SCALE), np.float32(SPECTRAL_VSCALE / (self.Ny * self.Nx)), self.opts,
             ptr(self.net.u), self.net.st)

    def loss(self, psi_c, target, B, grad=False):
        r0, c0, H, W = self.window
        call("namsa_spectral_loss", ptr(self.net.out), ptr(psi_c), ptr(target), B, self.Ny, self.Nx, r0, c0, H, W,
             np.float32(self.s), ptr(self.net.sums), ptr(self.net.gout) if grad else None, self.net.st)
        return self.net.sums[:2 * B].view(B, 2)

    def apply(self, psi_c, B, cbed):
        r0, c0, H, W = self.window
        call("namsa_spectral_apply", ptr(self.net.out), ptr(psi_c), B, self.Ny, self.Nx, r0, c0, H, W,
             np.float32(1.0), ptr(cbed), self.net.st)
        return cbed

    def correct(self, psi_c, psi_0, vhat, vidx, kmax, thick, cbed):
        """CBEDs [B][Ny][Nx] of corrected coarse exit spectra (kmax, thick: device [B] float32)."""
        B = psi_c.shape[0]
        self.pack(psi_c, psi_0, vhat, vidx, kmax, thick, B)
        if self.fused and self.net.fused_ok():
            self.net.forward_graph(B)
        else:
            self.net.forward(B)
        return self.apply(psi_c, B, cbed)


class SpectralPairPool:
    """Training pairs resident in HBM: per coarsening factor the coarse exit
    spectra, and shared by every factor the vacuum spectra, the fine CBEDs
    (targets), the thicknesses and an index into the potential spectra."""

    def __init__(self, factors=(8,)):
        self.factors = tuple(int(f) for f in factors)
        self.pc = {f: [] for f in self.factors}
        self.p0, self.f, self.th, self.vi, self.vh, self.km = [], [], [], [], [], []

    def add(self, mat, pos, thicknesses):
        """Pairs of the probes at pos at each thickness (a multiple of every
        factor's coarse slice)."""
        torch = _torch()
        dev = pos.device
        B = pos.shape[0]
        for T in thicknesses:
            S_f = int(round(T / mat.dz))
            p0 = torch.empty((B, mat.Ny, mat.Nx), dtype=torch.complex64, device=dev)
            f = torch.empty((B, mat.Ny, mat.Nx), dtype=torch.float32, device=dev)
            vh = torch.empty((mat.Ny, mat.Nx), dtype=torch.complex64, device=dev)
            mat.vacuum_spectra(pos, S_f * mat.dz, p0)
            mat.cbed(pos, S_f, None, f)
            mat.potential_spectrum(S_f, vh)
            for fac in self.factors:
                if S_f % fac:
                    raise NamsaError(f"thickness {T} is not a whole number of {fac * mat.dz} A slices")
                pc = torch.empty((B, mat.Ny, mat.Nx), dtype=torch.complex64, device=dev)
                mat.exit_spectra(pos, S_f // fac, fac, pc)
                self.pc[fac].extend(pc[b] for b in range(B))
            vi = len(self.vh)
            self.vh.append(vh)
            for b in range(B):
                self.p0.append(p0[b])
                self.f.append(f[b])
                self.th.append(float(S_f * mat.dz))
                self.vi.append(vi)
                self.km.append(float(mat.sim.kmax))

    def finalize(self):
        torch = _torch()
        self.PC = {fac: torch.stack(v) for fac, v in self.pc.items()}
        self.P0 = torch.stack(self.p0)
        self.F = torch.stack(self.f)
        self.VH = torch.stack(self.vh)
        dev = self.P0.device
        self.TH = torch.tensor(self.th, dtype=torch.float32, device=dev)
        self.VI = torch.tensor(self.vi, dtype=torch.int32, device=dev)
        self.KM = torch.tensor(self.km, dtype=torch.float32, device=dev)
        self.pc = self.p0 = self.f = self.vh = None
        return self

    def __len__(self):
        return len(self.th)

    def batch(self, idx, factor):
        torch = _torch()
        ii = torch.as_tensor(np.asarray(idx), device=self.P0.device)
        return (self.PC[factor].index_select(0, ii), self.P0.index_select(0, ii), self.F.index_select(0, ii),
                self.TH.index_select(0, ii), self.VI.index_select(0, ii), self.KM.index_select(0, ii))


def spectral_batch_loss(corr, pool, idx, grad=False):
    """Per-sample (sum (I - F)^2, sum F^2) of pool samples idx under the
    network (and with grad the gradient of the mean relative error in net.grad)."""
    B = len(idx)
    pc, p0, f, th, vi, km = pool.batch(idx, corr.factor)
    corr.pack(pc, p0, pool.VH, vi, km, th, B)
    corr.net.forward(B)
    sums = corr.loss(pc, f, B, grad=grad)
    if grad:
        corr.net.backward(B)
    return sums


def spectral_train(corr, pool, steps, batch=16, lr=2e-3, seed=0, log_every=0, log=print):
    """Adam with cosine decay on random minibatches; returns (step, mean rel-L2) log points."""
    rng = np.random.default_rng(seed)
    corr.net.ensure(batch)
    hist = []
    for it in range(steps):
        idx = rng.choice(len(pool), size=batch, replace=False)
        sums = spectral_batch_loss(corr, pool, idx, grad=True)
        corr.net.adam(lr * 0.5 * (1 + math.cos(math.pi * it / steps)))
        if log_every and (it % log_every == 0 or it == steps - 1):
            r = float((sums[:, 0] / sums[:, 1]).sqrt().mean())
            hist.append((it, r))
            log(f"step {it:5d}  CBED rel-L2 {r:.4f}")
    return hist


def spectral_evaluate(corr, pool, batch=16):
    """Per-sample CBED rel-L2 over the window, learned and uncorrected (o = 0)."""
    torch = _torch()
    corr.net.ensure(batch)
    learned, coarse = [], []
    for i0 in range(0, len(pool), batch):
        idx = list(range(i0, min(len(pool), i0 + batch)))
        B = len(idx)
        s = spectral_batch_loss(corr, pool, idx).cpu().numpy()
        learned.extend(np.sqrt(s[:, 0] / s[:, 1]))
        pc, _, f, _, _, _ = pool.batch(idx, corr.factor)
        call("namsa_fill", ptr(corr.net.out), 2 * B * corr.window[2] * corr.window[3], np.float32(0.0), corr.net.st)
        s0 = corr.loss(pc, f, B).cpu().numpy()
        coarse.extend(np.sqrt(s0[:, 0] / s0[:, 1]))
    return np.asarray(learned), np.asarray(coarse)


# ---------------------------------------------------------------------- training data
def random_crystal_sim(seed, grid=256, pix=0.125, kv=200.0, semi=0.025, thickness=(40.0, 400.0), device=None):
    """A seeded random crystal (cubic / tetragonal / hexagonal lattice, 1-8
    basis atoms, Z in [3, 83]) grid * pix wide and uniform in `thickness`
    thick, prepared as a drop-in simulation at grid^2 (the structures of the
    corrector's training pairs).  Returns (sim, rng) with the rng advanced
    past the structure, for the pair's remaining random choices."""
    from . import structures
    from .msa import MSAGPU

    rng = np.random.default_rng(int(seed))
    T = float(rng.uniform(*thickness))
    rec, _ = structures.random_crystal(rng, grid * pix, T + 0.5)
    sim = MSAGPU(kv, semi, rec, sampling=np.array([grid, grid]), debye_waller=False)
    sim.setup_device(0 if device is None else device)
    sim.calc_atomic_potentials()
    return sim, rng, T


def polycrystal_sim(seed, grid=2048, kv=300.0, semi=0.030, lateral=128.0, thickness=1000.0, device=None):
    """A seeded 4-grain Si / GaAs / Au polycrystal lateral x lateral x
    thickness A (the large-scale comparison's structure type) as a drop-in
    simulation at grid^2.  Returns (sim, rng, thickness) like random_crystal_sim."""
    from . import structures
    from .msa import MSAGPU

    rec, _ = structures.polycrystal(seed=int(seed), lateral=lateral, thickness=thickness)
    sim = MSAGPU(kv, semi, rec, sampling=np.array([grid, grid]), debye_waller=False)
    sim.setup_device(0 if device is None else device)
    sim.calc_atomic_potentials()
    return sim, np.random.default_rng(int(seed) + 7919), float(thickness)


def make_spectral_pools(seeds, factors=(8,), probes=4, grid=256, pix=0.125, kv=200.0, semi=0.025,
                        thickness=(40.0, 400.0), dz=1.0, log=None, builder=None):
    """One pair pool over every coarsening factor, from seeded structures.

    builder(seed) -> (sim, rng, T) makes a structure (default: a random
    crystal, random_crystal_sim with the other arguments).  Per structure:
    `probes` random positions within the central half of the cell and two
    thickness snapshots -- the whole structure, and a random thickness from
    thickness[0] up, both in whole multiples of 8 dz -- so 2 * probes pairs.
    The fine targets, vacuum spectra and potential spectra are shared by
    every factor; only the coarse exit spectra differ.
    """
    torch = _torch()
    if builder is None:
        builder = lambda sd: random_crystal_sim(sd, grid, pix, kv, semi, thickness)   # noqa: E731
    pool = SpectralPairPool(factors)
    unit = 8 * dz
    mats = None
    for n, seed in enumerate(seeds):
        mats = None
        sim, rng, T = builder(seed)
        L = float(min(sim.dims[:2]))
        pos = torch.from_numpy(rng.uniform(-L / 4, L / 4, (probes, 2)).astype(np.float32)).to(sim.device)
        mats = SpectralMaterial(sim, dz=dz, factors=factors)
        units = int(np.floor(mats.S_f * dz / unit))
        lo = min(units, int(np.ceil(thickness[0] / unit)))
        snaps = sorted({units, int(rng.integers(lo, units + 1))})
        pool.add(mats, pos, [u * unit for u in snaps])
        if log and (n % 100 == 0 or n == len(seeds) - 1 or len(seeds) <= 20):
            log(f"  structure {n + 1}/{len(seeds)} (seed {seed}): {len(sim.supercell_Z)} atoms, "
                f"snapshots {[u * unit for u in snaps]} A, {len(pool)} pairs")
        del sim
    return pool.finalize(), mats


DEFAULT_TAU = 0.25   # fall back to fine slicing where the network changes the CBED by more than 25 %


def run_spectral(corr, mat, pos, thickness, chunk=None, vhat=None, tau=DEFAULT_TAU, stats=None):
    """The learned coarse path: CBEDs [B][Ny][Nx] of probes at pos through
    thickness / (factor dz) coarse slices, each probe's exit spectrum
    corrected once by the network.

    Fallback: the network's residual estimate for a probe is the relative
    change it makes to the coarse CBED, ||I_learned - I_coarse|| / ||I_learned||
    over the window; probes above tau are recomputed by plain fine slicing
    (tau None: never).  chunk: probes per network call (bounds the
    activation memory); vhat: the potential spectrum over the thickness
    (computed when None); stats: a dict that receives the estimates and the
    fallback count."""
    torch = _torch()
    B = pos.shape[0]
    f = corr.factor
    S_f = int(round(thickness / mat.dz))
    if S_f % f or S_f > mat.S_f:
        raise NamsaError(f"thickness {thickness} is not a whole number of {f * mat.dz} A slices of the material")
    dev = pos.device
    shape = (mat.Ny, mat.Nx)
    if vhat is None:
        vhat = mat.potential_spectrum(S_f, torch.empty(shape, dtype=torch.complex64, device=dev))
    chunk = max(1, min(int(chunk or B), B))
    cb = torch.empty((B,) + shape, dtype=torch.float32, device=dev)
    km = torch.full((chunk,), float(mat.sim.kmax), dtype=torch.float32, device=dev)
    th = torch.full((chunk,), float(S_f * mat.dz), dtype=torch.float32, device=dev)
    pc = torch.empty((chunk,) + shape, dtype=torch.complex64, device=dev)
    p0 = torch.empty_like(pc)
    est = torch.empty((B,), dtype=torch.float64, device=dev) if tau is not None else None
    r0, c0, H, W = corr.window
    for s in range(0, B, chunk):
        b = min(chunk, B - s)
        mat.exit_spectra(pos[s:s + b], S_f // f, f, pc[:b])
        mat.vacuum_spectra(pos[s:s + b], S_f * mat.dz, p0[:b])
        corr.correct(pc[:b], p0[:b], vhat.view((1,) + shape), None, km[:b], th[:b], cb[s:s + b])
        if tau is not None:
            # with the correction zeroed the loss kernel's I is the coarse CBED; against the learned CBED
            # as target it gives sum (I_coarse - I_learned)^2 and sum I_learned^2 per probe
            net = corr.net
            call("namsa_fill", ptr(net.out), 2 * b * H * W, np.float32(0.0), net.st)
            call("namsa_spectral_loss", ptr(net.out), ptr(pc), ptr(cb[s]), b, mat.Ny, mat.Nx, r0, c0, H, W,
                 np.float32(1.0), ptr(net.sums), None, net.st)
            sums = net.sums[:2 * b].view(b, 2)
            est[s:s + b] = (sums[:, 0] / sums[:, 1].clamp_min(1e-300)).sqrt()
    nf = 0
    if tau is not None:
        e = est.cpu().numpy()
        bad = np.nonzero(~(e <= tau))[0]          # NaN estimates fall back too
        nf = len(bad)
        if nf:
            idx = torch.as_tensor(bad, device=dev)
            fine = torch.empty((nf,) + shape, dtype=torch.float32, device=dev)
            mat.cbed(pos.index_select(0, idx), S_f, None, fine)
            cb.index_copy_(0, idx, fine)
        if stats is not None:
            stats["estimate"] = e
    if stats is not None:
        stats["fallbacks"] = nf
    return cb
