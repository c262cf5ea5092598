"""GPU kernels of the coarse-slice corrector (MI355X): the BF16 MFMA GEMM
against float64 numpy products of the same bf16-rounded operands (FP32
accumulation: tolerance 2e-5 of the result norm), and the pointwise kernels
against closed forms."""
import sys

import numpy as np
import pytest

from namsa_amd import _lib
from namsa_amd._lib import call, ptr
from oracle import fno_oracle as F

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bf16(a):
    """Round-to-nearest-even float32 -> bfloat16, returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _gemm(torch, A, sA, B, sB, Cbuf, sC, M, N, K, batch=1, split=1, alpha=1.0, beta=0.0, bias=None,
          bias_mode=0, gelu=False, P=None):
    call("namsa_gemm_bf16", ptr(A), *sA, ptr(B), *sB, ptr(Cbuf), *sC, M, N, K, batch, split,
         np.float32(alpha), np.float32(beta), ptr(bias) if bias is not None else None, bias_mode,
         1 if gelu else 0, ptr(P) if P is not None else None, _lib.stream_handle())
    torch.cuda.synchronize()


def _view(flat, b, s, rows, cols):
    """Gather the (rows x cols) matrix of batch b from strides s = (sb, si, sj)."""
    i = np.arange(rows)[:, None]
    j = np.arange(cols)[None, :]
    return flat[b * s[0] + i * s[1] + j * s[2]]


@pytest.mark.parametrize("M,N,K,batch,a_t,b_t", [
    (64, 64, 32, 1, False, False),     # one tile, one k step
    (32, 32, 16, 1, False, False),     # a single MFMA block (checks the lane layout)
    (100, 77, 45, 3, False, False),    # ragged edges in every dimension, batched
    (70, 130, 96, 2, True, False),     # transposed A
    (70, 130, 96, 2, False, True),     # transposed B
    (33, 65, 200, 4, True, True),
])
def test_gemm_layouts(gpu, M, N, K, batch, a_t, b_t):
    import torch
    _say(f"gemm M={M} N={N} K={K} batch={batch} aT={a_t} bT={b_t}")
    rng = np.random.default_rng(M * 7 + N + K)
    sA = (M * K + 5, 1, M) if a_t else (M * K + 5, K, 1)
    sB = (K * N + 3, 1, K) if b_t else (K * N + 3, N, 1)
    sC = (M * N + 1, N, 1)
    A = rng.standard_normal(batch * sA[0]).astype(np.float32)
    B = rng.standard_normal(batch * sB[0]).astype(np.float32)
    Ad, Bd = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    Cd = torch.full((batch * sC[0],), 7.0, dtype=torch.float32, device="cuda")
    _gemm(torch, Ad, sA, Bd, sB, Cd, sC, M, N, K, batch=batch)
    C = Cd.cpu().numpy()
    Ab, Bb = bf16(A), bf16(B)
    for b in range(batch):
        ref = _view(Ab, b, sA, M, K) @ _view(Bb, b, sB, K, N)
        got = _view(C, b, sC, M, N)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 2e-5, (b, err)
    # the padding word between batches is untouched
    assert C[M * N] == 7.0


@pytest.mark.parametrize("M,N,K,batch,sA,sB,offA,offB", [
    (300, 32, 200, 2, (300 * 200, 200, 1), (200 * 32, 32, 1), 0, 0),    # 128x32 tiles, 16-byte loads of A and B
    (300, 17, 70, 2, (300 * 72, 72, 1), (70 * 20, 20, 1), 0, 0),        # ragged N and K with padded strides
    (32, 1000, 32, 2, (32 * 32, 32, 1), (32 * 1000, 1000, 1), 0, 0),    # 32x128 tiles (a pointwise layer)
    (7, 300, 64, 1, (7 * 64, 64, 1), (64 * 300, 300, 1), 0, 0),         # ragged M below one tile
    (128, 200, 96, 2, (128 * 96, 1, 128), (96 * 200, 200, 1), 0, 0),    # A unit stride along i
    (64, 96, 40, 1, (64 * 40, 40, 1), (40 * 96, 1, 40), 0, 0),          # B unit stride along k
    (300, 32, 200, 1, (300 * 200, 200, 1), (200 * 32, 32, 1), 1, 3),    # misaligned bases: element loads
    (32, 32, 170, 128, (32 * 170, 170, 1), (170 * 32, 32, 1), 0, 0),    # K split over four wavefronts (y DFT)
    (20, 17, 40, 70, (20 * 40, 40, 1), (40 * 17 + 1, 1, 40), 0, 0),     # ragged items, K not a multiple of 16
    (20, 17, 8, 70, (20 * 8, 8, 1), (8 * 17, 17, 1), 0, 0),             # K below one step
    (170, 32, 32, 96, (170 * 32, 1, 170), (32 * 32, 32, 1), 0, 0),      # thin-K batched items (inverse y DFT)
])
def test_gemm_tile_shapes_and_vector_loads(gpu, M, N, K, batch, sA, sB, offA, offB):
    import torch
    _say(f"gemm tiles M={M} N={N} K={K} batch={batch} offsets=({offA},{offB})")
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    sC = (M * N, N, 1)
    A = rng.standard_normal(offA + batch * sA[0]).astype(np.float32)
    B = rng.standard_normal(offB + batch * sB[0]).astype(np.float32)
    Ad, Bd = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    Cd = torch.zeros(batch * sC[0], dtype=torch.float32, device="cuda")
    call("namsa_gemm_bf16", C_ptr(Ad, offA), *sA, C_ptr(Bd, offB), *sB, ptr(Cd), *sC, M, N, K, batch, 1,
         np.float32(1.0), np.float32(0.0), None, 0, 0, None, _lib.stream_handle())
    torch.cuda.synchronize()
    C = Cd.cpu().numpy()
    Ab, Bb = bf16(A[offA:]), bf16(B[offB:])
    for b in range(batch):
        ref = _view(Ab, b, sA, M, K) @ _view(Bb, b, sB, K, N)
        err = np.linalg.norm(_view(C, b, sC, M, N) - ref) / np.linalg.norm(ref)
        assert err < 2e-5, (b, err)


def C_ptr(t, off):
    """Device pointer `off` floats into tensor t."""
    import ctypes
    return ctypes.c_void_p(t.data_ptr() + 4 * off)


def test_gemm_epilogue_bias_beta_gelu_and_split_k(gpu):
    import torch
    _say("gemm epilogues")
    rng = np.random.default_rng(5)
    M, N, K = 96, 80, 300
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32) * 0.1
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    bias_i = rng.standard_normal(M).astype(np.float32)
    bias_j = rng.standard_normal(N).astype(np.float32)
    Ad, Bd = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    ref = bf16(A) @ bf16(B)
    s = (0, K, 1), (0, N, 1), (0, N, 1)
    # alpha, beta, row bias, GELU with the pre-activation stored
    Cd = torch.from_numpy(C0.copy()).cuda()
    Pd = torch.zeros_like(Cd)
    _gemm(torch, Ad, s[0], Bd, s[1], Cd, s[2], M, N, K, alpha=0.5, beta=2.0,
          bias=torch.from_numpy(bias_i).cuda(), bias_mode=1, gelu=True, P=Pd)
    pre = 0.5 * ref + 2.0 * C0 + bias_i[:, None]
    np.testing.assert_allclose(Pd.cpu().numpy(), pre, rtol=0, atol=2e-5 * np.abs(pre).max())
    np.testing.assert_allclose(Cd.cpu().numpy(), F.gelu(pre), rtol=0, atol=2e-5 * np.abs(pre).max())
    # column bias
    Cd = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    _gemm(torch, Ad, s[0], Bd, s[1], Cd, s[2], M, N, K, bias=torch.from_numpy(bias_j).cuda(), bias_mode=2)
    np.testing.assert_allclose(Cd.cpu().numpy(), ref + bias_j[None], rtol=0, atol=2e-5 * np.abs(ref).max())
    # split-K adds onto C (atomics): 5 splits of a K=300 product onto C0
    Cd = torch.from_numpy(C0.copy()).cuda()
    _gemm(torch, Ad, s[0], Bd, s[1], Cd, s[2], M, N, K, split=5, alpha=-1.0)
    np.testing.assert_allclose(Cd.cpu().numpy(), C0 - ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    # the same epilogue on the batched small-item path (32 x 32 items) and on batched tiles
    for m, n, k in ((32, 24, 100), (90, 30, 48)):
        nb = 64
        a = rng.standard_normal((nb, m, k)).astype(np.float32)
        bm = rng.standard_normal((nb, k, n)).astype(np.float32) * 0.1
        c0 = rng.standard_normal((nb, m, n)).astype(np.float32)
        bi = rng.standard_normal(m).astype(np.float32)
        Cd = torch.from_numpy(c0.copy()).cuda()
        Pd = torch.zeros_like(Cd)
        _gemm(torch, torch.from_numpy(a).cuda(), (m * k, k, 1), torch.from_numpy(bm).cuda(), (k * n, n, 1), Cd,
              (m * n, n, 1), m, n, k, batch=nb, alpha=0.5, beta=2.0, bias=torch.from_numpy(bi).cuda(), bias_mode=1,
              gelu=True, P=Pd)
        pre = 0.5 * np.einsum("bik,bkj->bij", bf16(a).reshape(nb, m, k), bf16(bm).reshape(nb, k, n)) + 2.0 * c0 \
            + bi[None, :, None]
        np.testing.assert_allclose(Pd.cpu().numpy(), pre, rtol=0, atol=2e-5 * np.abs(pre).max())
        np.testing.assert_allclose(Cd.cpu().numpy(), F.gelu(pre), rtol=0, atol=2e-5 * np.abs(pre).max())


def test_gemm_rejects_bad_arguments(gpu):
    import torch
    x = torch.zeros(16, device="cuda")
    lib = _lib.load()
    st = _lib.stream_handle()
    rc = lib.namsa_gemm_bf16(ptr(x), 0, 1, 1, ptr(x), 0, 1, 1, ptr(x), 0, 1, 1, 4, 4, 4, 1, 2,
                             1.0, 0.0, ptr(x), 1, 0, None, st)
    assert rc != 0 and b"split-K" in lib.namsa_last_error()
    rc = lib.namsa_gemm_bf16(ptr(x), 0, 1, 1, ptr(x), 0, 1, 1, ptr(x), 0, 1, 1, 4, 4, 4, 1, 1,
                             1.0, 0.0, None, 0, 1, None, st)
    assert rc != 0


def test_pointwise_kernels(gpu):
    import torch
    _say("fill / gelu backward / adam")
    rng = np.random.default_rng(9)
    n = 5000
    x = torch.empty(n, device="cuda")
    call("namsa_fill", ptr(x), n, np.float32(2.5), _lib.stream_handle())
    assert torch.all(x == 2.5)
    pre = rng.standard_normal(n).astype(np.float32) * 3
    g = rng.standard_normal(n).astype(np.float32)
    gd = torch.from_numpy(g.copy()).cuda()
    pre_d = torch.from_numpy(pre).cuda()
    call("namsa_gelu_backward", ptr(gd), ptr(pre_d), n, _lib.stream_handle())
    # fp32 tanh: absolute error ~1e-6 where gelu' passes through zero near x = -0.75 and x < -3
    np.testing.assert_allclose(gd.cpu().numpy(), g * F.gelu_grad(pre.astype(np.float64)), rtol=2e-5, atol=1e-5)
    # Adam, two steps, against the textbook update with decoupled decay
    p0 = rng.standard_normal(n).astype(np.float32)
    grads = [rng.standard_normal(n).astype(np.float32) for _ in range(2)]
    pd, md, vd = torch.from_numpy(p0.copy()).cuda(), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    p, m, v = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    for step, gr in enumerate(grads, 1):
        gr_d = torch.from_numpy(gr).cuda()
        call("namsa_adam", ptr(pd), ptr(gr_d), ptr(md), ptr(vd), n, np.float32(lr),
             np.float32(b1), np.float32(b2), np.float32(eps), step, np.float32(wd), _lib.stream_handle())
        m = b1 * m + (1 - b1) * gr
        v = b2 * v + (1 - b2) * gr.astype(np.float64) ** 2
        p = p - lr * ((m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps) + wd * p)
    np.testing.assert_allclose(pd.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_mode_bookkeeping(gpu):
    import torch
    _say("modes combine / expand")
    rng = np.random.default_rng(4)
    C, B, m = 3, 2, 4
    m2 = 2 * m
    P = rng.standard_normal((C * B, m2, m2)).astype(np.float32)
    Xs = torch.empty((m * m, 2 * C, B), device="cuda")
    call("namsa_modes_combine", ptr(torch.from_numpy(P).cuda()), ptr(Xs), C, B, m, _lib.stream_handle())
    got = Xs.cpu().numpy()
    Pc = P.reshape(C, B, 2, m, 2, m)                     # [c][b][r][p][t][q]
    Xr = Pc[:, :, 0, :, 0, :] + Pc[:, :, 1, :, 1, :]      # [c][b][p][q]
    Xi = Pc[:, :, 0, :, 1, :] - Pc[:, :, 1, :, 0, :]
    ref = np.concatenate([Xr, Xi], 0).transpose(2, 3, 0, 1).reshape(m * m, 2 * C, B)
    np.testing.assert_allclose(got, ref, rtol=1e-6)
    Ys = rng.standard_normal((m * m, 2 * C, B)).astype(np.float32)
    E = torch.empty((C * B, m2, m2), device="cuda")
    call("namsa_modes_expand", ptr(torch.from_numpy(Ys).cuda()), ptr(E), C, B, m, _lib.stream_handle())
    e = E.cpu().numpy().reshape(C, B, 2, m, 2, m)
    Y = Ys.reshape(m, m, 2, C, B).transpose(2, 3, 4, 0, 1)   # [r][o][b][p][q]
    np.testing.assert_array_equal(e[:, :, 0, :, 0, :], Y[0])
    np.testing.assert_array_equal(e[:, :, 1, :, 0, :], -Y[1])
    np.testing.assert_array_equal(e[:, :, 0, :, 1, :], Y[1])
    np.testing.assert_array_equal(e[:, :, 1, :, 1, :], Y[0])


def test_row_fft_round_trip_and_exit_wave(gpu):
    """Inverse row transform (scale 1/Nx) of the slice state is the exit wave
    the EXIT mode writes (times its out_scale), and the forward transform
    takes it back to the state."""
    import torch
    _say("row transforms vs EXIT mode")
    from namsa_amd.msa import MSAGPU
    from tests.parity_util import random_block
    rec = random_block(3, 24.0, 10.0, 300)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    sim.build_probe()
    sim.probe_positions = np.array([[0.5, -1.0], [2.0, 3.0]])
    sim.num_probes = 2
    sim.plan_simulation()
    sim.multislice(output="exit_wave")
    exit_ref = sim.probes_device.clone()                     # scaled by 1/sqrt(norm)
    Ny, Nx = sim._grid
    S = int(sim.num_slices)
    pos = torch.from_numpy(sim.probe_positions.astype(np.float32)).cuda()
    ws = torch.empty((2, Ny, Nx), dtype=torch.complex64, device="cuda")
    st = _lib.stream_handle()
    call("namsa_probe_init", ptr(sim.psi_k_device), ptr(pos), 2, Ny, Nx, np.float32(sim.kmax), ptr(ws), st)
    for i in range(S):
        call("namsa_slice_rows", ptr(ws), 2, Ny, Nx, ptr(sim.potential_slices_device[i]), 0, Nx, st)
        call("namsa_slice_cols", ptr(ws), 2, Ny, Nx, ptr(sim.pm_device), 0, Nx, st)
    state = ws.clone()
    call("namsa_row_fft", ptr(ws), 2, Ny, Nx, 1, np.float32(1.0 / Nx), st)
    torch.cuda.synchronize()
    x = ws.cpu().numpy().astype(np.complex128)
    ref = exit_ref.cpu().numpy() * np.sqrt(sim.normalization) / Nx
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-5
    call("namsa_row_fft", ptr(ws), 2, Ny, Nx, 0, np.float32(1.0), st)
    torch.cuda.synchronize()
    back = ws.cpu().numpy()
    s0 = state.cpu().numpy()
    assert np.linalg.norm(back - s0) / np.linalg.norm(s0) < 1e-5
    lib = _lib.load()
    assert lib.namsa_row_fft(ptr(ws), 2, Ny, Nx, 0, 0.5, st) != 0
