"""Every compiled row/column schedule, at every transform length, against the
FP64 oracle: exit waves and CBED after three slices with random unit-modulus
slices and a random propagator that is zero outside a column band not aligned
to any strip width, run in probe groups of 2 (rel-L2 < 1e-5, FP32 rounding of
eight transforms)."""
import sys

import numpy as np
import pytest

import namsa_amd._lib as L
from namsa_amd import autotune
from oracle import msa_oracle as O
from tests.parity_util import rel_l2

pytestmark = pytest.mark.gpu
SIZES = [128, 192, 256, 384, 512, 768, 1024, 2048]


def _case(N, seed):
    rng = np.random.default_rng(seed)
    B, S = 3, 3
    psi_k = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N
    pos = rng.uniform(-5, 5, (B, 2)).astype(np.float32)
    t = np.exp(1j * rng.uniform(0, 2 * np.pi, (S, N, N)))
    band = (N // 5 + 3, N - N // 4 - 1)
    pm = np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N))) * (rng.random((N, N)) < 0.8)
    pm[:, :band[0]] = 0
    pm[:, band[1]:] = 0
    kmax = np.float32(6.5)
    exit_ref = O.multislice(O.probe_stack(psi_k, pos, kmax), t, pm)
    cbed_ref = O.readout(exit_ref, 1.0)
    return psi_k, pos, t, pm, band, kmax, exit_ref, cbed_ref


def _run(torch, dev, N, psi_k, pos, t, pm, band, kmax, mode):
    B, S = len(pos), len(t)
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)
    pk, tt, pmd = c64(psi_k), c64(t), c64(pm)
    p = torch.from_numpy(pos).to(dev)
    ws = torch.empty((B, N, N), dtype=torch.complex64, device=dev)
    cb = torch.empty((B, N, N), dtype=torch.float32, device=dev)
    L.call("namsa_multislice", L.ptr(pk), L.ptr(p), B, N, N, kmax, L.ptr(tt), S, L.ptr(pmd), band[0], band[1],
           2, mode, np.float32(1.0), L.ptr(ws), L.ptr(cb), None, 0, None, None, L.stream_handle())
    return (ws if mode == L.OUT_EXIT else cb).cpu().numpy()


@pytest.mark.parametrize("N", SIZES)
def test_every_schedule_matches_oracle(gpu, N):
    import torch

    psi_k, pos, t, pm, band, kmax, exit_ref, cbed_ref = _case(N, N + 1)
    keep = (autotune.selected(N, 0), autotune.selected(N, 1))
    try:
        for pass_ in (0, 1):
            for v, name in enumerate(autotune.schedules(N, pass_)):
                autotune.select(N, 0, v if pass_ == 0 else 0)
                autotune.select(N, 1, v if pass_ == 1 else 0)
                print(f"N={N} pass={pass_} schedule={name}", flush=True)
                sys.stdout.flush()
                ex = _run(torch, gpu, N, psi_k, pos, t, pm, band, kmax, L.OUT_EXIT)
                cb = _run(torch, gpu, N, psi_k, pos, t, pm, band, kmax, L.OUT_CBED)
                e1, e2 = rel_l2(ex, exit_ref), rel_l2(cb, cbed_ref)
                print(f"   exit relL2={e1:.2e} cbed relL2={e2:.2e}", flush=True)
                assert e1 < 1e-5 and e2 < 1e-5, (name, e1, e2)
    finally:
        autotune.select(N, 0, keep[0])
        autotune.select(N, 1, keep[1])


@pytest.mark.parametrize("N,reg,lds", [(512, "e8.r888.c8.pg.sb.w6", "e8.r888.c8.lr.pg.sb.w6"),
                                        (1024, "e16.r16.8.8.c8.pg.sb.w4", "e16.r16.8.8.c8.lr.pg.sb.w4"),
                                        (2048, "e16.r16.16.8.c4.pg.sb.w4", "e16.r16.16.8.c4.lr.pg.sb.w4")])
def test_lds_root_twiddles_match_register_twiddles(gpu, N, reg, lds):
    """The exact twiddles formed from the LDS quarter table of roots (`.lr`
    schedules) are the same float pairs as the register tables hold (checked
    on the host for every exponent at 2048), so the two transforms differ only
    where the compiler pairs a product differently inside a fused
    multiply-add: float32 round-off, not a twiddle error (propagator read at
    use in both)."""
    import torch

    psi_k, pos, t, pm, band, kmax, exit_ref, _ = _case(N, 7)
    names = autotune.schedules(N, 1)
    keep = autotune.selected(N, 1)
    out = {}
    try:
        for name in (reg, lds):
            autotune.select(N, 1, names.index(name))
            print(f"N={N} column schedule={name}", flush=True)
            out[name] = _run(torch, gpu, N, psi_k, pos, t, pm, band, kmax, L.OUT_EXIT)
    finally:
        autotune.select(N, 1, keep)
    a, b = out.values()
    d = rel_l2(b, a)
    print(f"   register vs LDS-root twiddles: relL2 {d:.2e}; vs oracle {rel_l2(a, exit_ref):.2e} {rel_l2(b, exit_ref):.2e}")
    assert d < 1e-6


def _tile(a):
    """[B][Ny][Nx] -> the same planes in 2 x 4 tiles: element (y, x) at
    ((y//2)(Nx//4) + x//4)*8 + (y%2)*4 + x%4 (namsa_slice_rows_tiled)."""
    B, Ny, Nx = a.shape
    return np.ascontiguousarray(a.reshape(B, Ny // 2, 2, Nx // 4, 4).transpose(0, 1, 3, 2, 4)).reshape(B, Ny, Nx)


def _untile(a):
    B, Ny, Nx = a.shape
    return np.ascontiguousarray(a.reshape(B, Ny // 2, Nx // 4, 2, 4).transpose(0, 1, 3, 2, 4)).reshape(B, Ny, Nx)


def test_tile_helpers_restate_the_offset():
    Ny, Nx = 6, 12
    a = np.arange(Ny * Nx).reshape(1, Ny, Nx)
    flat = _tile(a).reshape(-1)
    for y in range(Ny):
        for x in range(Nx):
            assert flat[(y // 2) * 2 * Nx + (x // 4) * 8 + (y % 2) * 4 + x % 4] == a[0, y, x]
    assert np.array_equal(_untile(_tile(a)), a)


@pytest.mark.parametrize("precise", [False, True])
def test_tiled_slice_passes_match_row_major(gpu, precise):
    """The slice step on the 2 x 4 tiled wave (the layout namsa_multislice's
    CBED loop keeps at 2048^2) computes the same float operations as on the
    row-major wave, only at other addresses: bit-identical after untiling,
    for the default schedules and their thick-run partners."""
    import torch

    N, B = 2048, 2
    lib = L.load()
    assert lib.namsa_wave_tiled(N, N) == 1
    rng = np.random.default_rng(3)
    psi = (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(np.complex64)
    t = np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N))).astype(np.complex64)
    pm = (np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N))) * (rng.random((N, N)) < 0.9)).astype(np.complex64)
    lo, hi = autotune.default_band(N)
    pm[:, :lo] = 0
    pm[:, hi:] = 0
    keep = (autotune.selected(N, 0), autotune.selected(N, 1))
    try:
        if precise:
            autotune.select(N, 0, lib.namsa_sched_precise(N, 0, keep[0]))
            autotune.select(N, 1, lib.namsa_sched_precise(N, 1, keep[1]))
        names = (autotune.schedules(N, 0)[autotune.selected(N, 0)], autotune.schedules(N, 1)[autotune.selected(N, 1)])
        print(f"N={N} B={B} schedules {names}", flush=True)
        td = torch.from_numpy(t).to(gpu)
        st = L.stream_handle()
        out = []
        for sfx, a, p in (("", psi, pm), ("_tiled", _tile(psi), _tile(pm[None])[0])):
            w = torch.from_numpy(a).to(gpu)
            pmd = torch.from_numpy(p).to(gpu)
            for band in ((0, N), (lo, hi)):
                L.call("namsa_slice_rows" + sfx, L.ptr(w), B, N, N, L.ptr(td), band[0], band[1], st)
                L.call("namsa_slice_cols" + sfx, L.ptr(w), B, N, N, L.ptr(pmd), lo, hi, st)
            out.append(w.cpu().numpy())
    finally:
        autotune.select(N, 0, keep[0])
        autotune.select(N, 1, keep[1])
    rm, tl = out[0], _untile(out[1])
    assert np.isfinite(rm).all()
    assert np.array_equal(rm.view(np.uint64), tl.view(np.uint64))


def test_tile_plane_and_tiled_propagator(gpu):
    """namsa_tile_plane writes the 2 x 4 tile layout, and 2048^2 CBED and
    exit-wave runs given the tiled propagator (NAMSA_PM_TILED) read the same
    values as runs given the row-major propagator: identical bits."""
    import torch

    N, B, S = 2048, 2, 3
    rng = np.random.default_rng(11)
    pm = (np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N))) * (rng.random((N, N)) < 0.9)).astype(np.complex64)
    lo, hi = autotune.default_band(N)
    pm[:, :lo] = 0
    pm[:, hi:] = 0
    src = torch.from_numpy(pm).to(gpu)
    dst = torch.empty_like(src)
    st = L.stream_handle()
    print(f"tile_plane {N}x{N}", flush=True)
    L.call("namsa_tile_plane", L.ptr(src), L.ptr(dst), N, N, st)
    assert np.array_equal(dst.cpu().numpy(), _tile(pm[None])[0])
    psi_k = torch.from_numpy(((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N)
                             .astype(np.complex64)).to(gpu)
    pos = torch.from_numpy(rng.uniform(-5, 5, (B, 2)).astype(np.float32)).to(gpu)
    t = torch.from_numpy(np.exp(1j * rng.uniform(0, 2 * np.pi, (S, N, N))).astype(np.complex64)).to(gpu)
    for mode in (L.OUT_CBED, L.OUT_EXIT):
        out = []
        for p, flag in ((src, 0), (dst, L.PM_TILED)):
            print(f"multislice {N}x{N} B={B} S={S} mode={mode} flags={flag}", flush=True)
            ws = torch.empty((B, N, N), dtype=torch.complex64, device=gpu)
            cb = torch.empty((B, N, N), dtype=torch.float32, device=gpu)
            L.call("namsa_multislice", L.ptr(psi_k), L.ptr(pos), B, N, N, np.float32(6.5), L.ptr(t), S, L.ptr(p),
                   lo, hi, B, mode | flag, np.float32(1.0), L.ptr(ws), L.ptr(cb), None, 0, None, None, st)
            out.append((cb if mode == L.OUT_CBED else ws).cpu().numpy())
        assert np.isfinite(out[0]).all() and np.abs(out[0]).max() > 0
        assert np.array_equal(out[0], out[1])
