"""The allocation-free multislice entry point against the allocating one (MI355X).

namsa_multislice_noalloc keeps the wave state and the detector partial sums
in the library's fixed per-device scratch array.  Its CBEDs and detector
signals must equal namsa_multislice's bit for bit when both push the same
probe groups through the same schedules; two calls in a row (the second with
fewer probes and other positions) check that nothing left in the shared
array from the first call is read by the second.
"""
import sys

import numpy as np
import pytest

import namsa_amd._lib as L
from namsa_amd import DEFAULT_PROBE, MSAGPU
from tests.parity_util import random_block

pytestmark = pytest.mark.gpu
SCRATCH_BYTES = 64 << 20   # NAMSA_SCRATCH_BYTES in include/namsa_hip.h


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _capacity(N, n_det):
    return SCRATCH_BYTES // (8 * N * N + 8 * (N // 4) * n_det)


@pytest.fixture(scope="module")
def sim512(gpu):
    rec = random_block(71, 64.0, 24.0, 2500)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([512, 512]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.build_propagator(1 / 3)
    return sim


def _run(sim, pos, group, det, noalloc):
    import torch

    dev = sim.device
    N = 512
    B = pos.shape[0]
    nDet = det.shape[0]
    st = L.stream_handle()
    pos_d = torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float32)).to(dev)
    cbed = torch.full((B, N, N), np.nan, dtype=torch.float32, device=dev)
    det_out = torch.full((B, nDet), np.nan, dtype=torch.float64, device=dev)
    args = (L.ptr(sim.psi_k_device), L.ptr(pos_d), B, N, N, np.float32(sim.kmax), L.ptr(sim.potential_slices_device),
            int(sim.num_slices), L.ptr(sim.pm_device), sim.band[0], sim.band[1])
    scale = np.float32(1.0 / sim.normalization)
    if noalloc:
        L.call("namsa_multislice_noalloc", *args, group, scale, L.ptr(cbed), L.ptr(det), nDet, L.ptr(det_out), st)
    else:
        G = min(group, B)
        psi_ws = torch.empty((G, N, N), dtype=torch.complex64, device=dev)
        det_ws = torch.empty(int(L.load().namsa_det_workspace_doubles(G, N, N, nDet)), dtype=torch.float64,
                             device=dev)
        L.call("namsa_multislice", *args, G, L.OUT_CBED, scale, L.ptr(psi_ws), L.ptr(cbed), L.ptr(det), nDet,
               L.ptr(det_ws), L.ptr(det_out), st)
    torch.cuda.synchronize()
    return cbed.cpu().numpy(), det_out.cpu().numpy()


def test_noalloc_matches_multislice_two_calls(sim512):
    import torch

    sim = sim512
    det = torch.from_numpy(np.stack([sim.detector_mask(0.0, 0.025, mode="physical"),
                                     sim.detector_mask(0.040, 0.060, mode="physical")])).to(sim.device)
    cap = _capacity(512, 2)
    assert cap == 31
    rng = np.random.default_rng(5)
    pos1 = rng.uniform(-20, 20, (40, 2))            # two groups: 31 + 9
    pos2 = rng.uniform(-20, 20, (7, 2))             # one smaller group, other positions
    for k, (pos, group) in enumerate(((pos1, 0), (pos2, 0))):
        B = pos.shape[0]
        G = min(B, cap)
        _say(f"noalloc call {k + 1}: 512^2, {int(sim.num_slices)} slices, B={B}, group={G}, 2 detectors")
        got_c, got_d = _run(sim, pos, group, det, noalloc=True)
        ref_c, ref_d = _run(sim, pos, G, det, noalloc=False)
        assert np.isfinite(got_c).all() and np.isfinite(got_d).all()
        np.testing.assert_array_equal(got_c, ref_c)
        np.testing.assert_array_equal(got_d, ref_d)
        tot = got_c.reshape(B, -1).sum(1)
        assert (tot > 0.9).all() and (tot < 1.0 + 1e-5).all()


def test_noalloc_explicit_group_within_capacity(sim512):
    import torch

    sim = sim512
    det = torch.ones((1, 512, 512), dtype=torch.float32, device=sim.device)
    pos = np.random.default_rng(6).uniform(-20, 20, (12, 2))
    _say("noalloc explicit group 5, B=12")
    got_c, got_d = _run(sim, pos, 5, det, noalloc=True)
    ref_c, ref_d = _run(sim, pos, 5, det, noalloc=False)
    np.testing.assert_array_equal(got_c, ref_c)
    np.testing.assert_array_equal(got_d, ref_d)


def test_noalloc_rejects_groups_beyond_scratch(gpu):
    lib = L.load()
    # 512^2 with two detectors holds 31 probes; an explicit 32 is refused before any launch
    rc = lib.namsa_multislice_noalloc(None, None, 64, 512, 512, np.float32(8.0), None, 10, None, 0, 512, 32,
                                      np.float32(1.0), None, None, 2, None, None)
    assert rc == -1
    assert b"exceeds" in lib.namsa_last_error()
    # 2048^2: one probe plus two detectors' partials fit, two do not
    rc = lib.namsa_multislice_noalloc(None, None, 4, 2048, 2048, np.float32(8.0), None, 10, None, 0, 2048, 2,
                                      np.float32(1.0), None, None, 2, None, None)
    assert rc == -1
    assert b"exceeds" in lib.namsa_last_error()


def test_noalloc_tiled_2048_matches_multislice(gpu):
    """At 2048^2 both entry points keep the wave in 2 x 4 tiles (one probe per
    group in the 64 MiB scratch); namsa_multislice given the tiled propagator
    (NAMSA_PM_TILED) and the allocation-free call given the row-major one
    read the same values: identical CBEDs and detector signals."""
    import torch

    N, B, S = 2048, 2, 3
    assert _capacity(N, 1) == 1
    rng = np.random.default_rng(29)
    st = L.stream_handle()
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(gpu)
    psi_k = c64((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N)
    pos = torch.from_numpy(rng.uniform(-5, 5, (B, 2)).astype(np.float32)).to(gpu)
    t = c64(np.exp(1j * rng.uniform(0, 2 * np.pi, (S, N, N))))
    lo, hi = N // 6, N - N // 6
    pm = np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N)))
    pm[:, :lo] = 0
    pm[:, hi:] = 0
    pm = c64(pm)
    pm_t = torch.empty_like(pm)
    L.call("namsa_tile_plane", L.ptr(pm), L.ptr(pm_t), N, N, st)
    det = torch.from_numpy(rng.random((1, N, N)).astype(np.float32)).to(gpu)
    out = []
    for noalloc in (False, True):
        _say(f"2048^2 B={B} S={S} noalloc={noalloc}")
        cbed = torch.full((B, N, N), np.nan, dtype=torch.float32, device=gpu)
        det_out = torch.full((B, 1), np.nan, dtype=torch.float64, device=gpu)
        if noalloc:
            L.call("namsa_multislice_noalloc", L.ptr(psi_k), L.ptr(pos), B, N, N, np.float32(6.5), L.ptr(t), S,
                   L.ptr(pm), lo, hi, 0, np.float32(1.0), L.ptr(cbed), L.ptr(det), 1, L.ptr(det_out), st)
        else:
            ws = torch.empty((1, N, N), dtype=torch.complex64, device=gpu)
            det_ws = torch.empty(int(L.load().namsa_det_workspace_doubles(1, N, N, 1)), dtype=torch.float64,
                                 device=gpu)
            L.call("namsa_multislice", L.ptr(psi_k), L.ptr(pos), B, N, N, np.float32(6.5), L.ptr(t), S, L.ptr(pm_t),
                   lo, hi, 1, L.OUT_CBED | L.PM_TILED, np.float32(1.0), L.ptr(ws), L.ptr(cbed), L.ptr(det), 1,
                   L.ptr(det_ws), L.ptr(det_out), st)
        torch.cuda.synchronize()
        out.append((cbed.cpu().numpy(), det_out.cpu().numpy()))
    (c0, d0), (c1, d1) = out
    assert np.isfinite(c0).all() and c0.max() > 0
    assert np.array_equal(c0, c1) and np.array_equal(d0, d1)
