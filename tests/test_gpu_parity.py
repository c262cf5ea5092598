"""HIP path vs the FP64 oracle on the same seeded inputs (MI355X).

Tolerances (stated per test): transforms and pointwise kernels to FP32
rounding; CBED intensities to 1e-4 relative L2 (the plain-method parity bar);
discrete decisions (bandwidth mask, aperture) bit-exact.
"""
import sys

import numpy as np
import pytest

import namsa_amd._lib as L
from namsa_amd import MSAGPU, structures
from oracle import msa_oracle as O
from tests.parity_util import (KIRK, oracle_pm, oracle_probe, oracle_slices, random_block,
                               rel_l2)

pytestmark = pytest.mark.gpu
SIZES = [128, 192, 256, 384, 512, 768, 1024, 2048]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev_c64(torch, arr, dev):
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex64)).to(dev)


@pytest.mark.parametrize("N", SIZES)
def test_fft2d_matches_numpy(gpu, N):
    import torch

    _say(f"fft2d N={N}")
    rng = np.random.default_rng(N)
    B = 3 if N <= 1024 else 2
    x = (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(np.complex64)
    for inverse in (0, 1):
        d = _dev_c64(torch, x, gpu)
        L.call("namsa_fft2d", L.ptr(d), B, N, N, inverse, L.stream_handle())
        got = d.cpu().numpy()
        x64 = x.astype(np.complex128)
        ref = np.fft.ifft2(x64) * N * N if inverse else np.fft.fft2(x64)
        err = rel_l2(got, ref)
        _say(f"  inverse={inverse} relL2={err:.2e}")
        assert err < 3e-6   # FP32 rounding of an N^2-point transform


def test_fft2d_rectangular(gpu):
    import torch

    rng = np.random.default_rng(1)
    x = (rng.standard_normal((2, 256, 512)) + 1j * rng.standard_normal((2, 256, 512))).astype(np.complex64)
    d = _dev_c64(torch, x, gpu)
    L.call("namsa_fft2d", L.ptr(d), 2, 256, 512, 0, L.stream_handle())
    assert rel_l2(d.cpu().numpy(), np.fft.fft2(x.astype(np.complex128))) < 3e-6


@pytest.mark.parametrize("pix", [(0.125, 0.125), (0.15, 0.148), (0.0625, 0.0625)])
def test_atomic_tables_match_oracle(gpu, pix):
    import torch

    _say(f"atomic tables pix={pix}")
    Zs = np.array([3, 6, 14, 29, 79, 83])
    Py, Px = O.table_shape(np.asarray(pix))
    params = torch.from_numpy(np.ascontiguousarray(KIRK[Zs - 1])).to(gpu)
    out = torch.empty((len(Zs), Py, Px), dtype=torch.float32, device=gpu)
    L.call("namsa_atomic_tables", L.ptr(params), len(Zs), Py, Px, 8.0, 2, L.ptr(out), L.stream_handle())
    got = out.cpu().numpy()
    for i, Z in enumerate(Zs):
        ref = O.atomic_table(KIRK[Z - 1], np.asarray(pix))
        # float32 storage of FP64 values: a few ulp of the table maximum
        np.testing.assert_allclose(got[i], ref, rtol=0, atol=4e-7 * ref.max())


def test_slices_match_oracle_small(gpu):
    _say("slices small 256^2")
    rec = random_block(11, 32.0, 12.0, 400)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    t_ref, tables = oracle_slices(sim, 2.0)
    for z, tab in zip(sim.unique_Z, tables):
        np.testing.assert_allclose(sim.cached_pots[int(z)], tab, rtol=0, atol=4e-7 * tab.max())
    got = sim.potential_slices
    assert got.shape == t_ref.shape
    # phase error = sigma * (float32 sum error of V) + sincosf rounding
    assert np.abs(got - t_ref).max() < 2e-5


def test_slices_odd_table_and_edges(gpu):
    _say("slices odd table 192^2")
    rec = random_block(12, 28.7, 6.0, 150)
    sim = MSAGPU(100, 0.02, rec, sampling=np.array([192, 192]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    assert sim.table_shape == (53, 53)   # floor(8 / (28.7/192)): odd, rint half-to-even placement
    sim.build_potential_slices(1.5)
    t_ref, _ = oracle_slices(sim, 1.5)
    assert np.abs(sim.potential_slices - t_ref).max() < 2e-5


def test_probe_and_propagator_match_oracle(gpu):
    _say("probe / propagator 512^2")
    rec = structures.random_amorphous_crystal(seed=2)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([512, 512]), debye_waller=False)
    sim.setup_device(0)
    sim.slice_t = 2.0
    for smooth in (True, False):
        sim.build_probe({**sim_probe(), "smooth_apert": smooth})
        psi_k, A = oracle_probe(sim, smooth=smooth)
        np.testing.assert_allclose(sim.apert, A, rtol=0, atol=2e-7)
        np.testing.assert_allclose(sim.psi_k, psi_k, rtol=0, atol=2e-7)
        psi, norm = O.probe_real(psi_k)
        assert sim.normalization == pytest.approx(norm, rel=1e-6)
        assert rel_l2(sim.psi, psi / np.sqrt(norm)) < 2e-6
    sim.build_propagator(1 / 3)
    pm = oracle_pm(sim, 2.0, 1 / 3)
    np.testing.assert_array_equal(sim.mask_device.cpu().numpy(), np.abs(pm) > 0)   # bit-exact mask
    np.testing.assert_allclose(sim.pm_device.cpu().numpy(), pm, rtol=0, atol=2e-7)


def test_probe_with_aberrations(gpu):
    rec = random_block(3, 40.0, 10.0, 100)
    sim = MSAGPU(300, 0.02, rec, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    pd = {**sim_probe(), "aberration_dict": {"C1": 50.0, "C3": 1.2e7, "C5": 3e9}, "scherzer": False}
    sim.build_probe(pd)
    psi_k, _ = oracle_probe(sim, C=(50.0, 1.2e7, 3e9))
    np.testing.assert_allclose(sim.psi_k, psi_k, rtol=0, atol=3e-6)


def sim_probe():
    from namsa_amd import DEFAULT_PROBE

    return dict(DEFAULT_PROBE)


def _run_sim(sim, dz, positions, bandwidth=1 / 3, **kw):
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(dz)
    sim.build_probe(sim_probe())
    sim.probe_positions = np.asarray(positions, dtype=np.float64)
    sim.num_probes = np.int32(len(positions))
    sim.grid_steps = np.array([len(positions), 1])
    sim.plan_simulation()
    sim.multislice(bandwidth=bandwidth, **kw)


@pytest.mark.parametrize("N,S_thick", [(256, 20.0), (512, 16.0), (384, 12.0)])
def test_cbed_parity(gpu, N, S_thick):
    _say(f"cbed parity N={N}")
    rec = random_block(20 + N, 24.0, S_thick, int(24 * 24 * S_thick * 0.06))
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([N, N]), debye_waller=False)
    pos = np.array([[0.0, 0.0], [3.3, -2.1], [-5.25, 7.7], [0.01, 0.02], [6.0, 6.0]])
    _run_sim(sim, 2.0, pos)
    t_ref, _ = oracle_slices(sim, 2.0)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 2.0, 1 / 3))
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(len(pos))]
    _say(f"  S={sim.num_slices} CBED relL2 per probe {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-4


@pytest.mark.parametrize("Ny,Nx", [(256, 512), (384, 256)])
def test_cbed_parity_rectangular(gpu, Ny, Nx):
    # sampling = [Ny, Nx] with Ny != Nx: per-axis pixel sizes, one scalar kmax
    # (msa.py:114,121-122), the band on the x columns, different row and
    # column transform lengths
    _say(f"cbed parity rectangular Ny={Ny} Nx={Nx}")
    rec = random_block(7 + Ny + Nx, 24.0, 12.0, int(24 * 24 * 12.0 * 0.06))
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([Ny, Nx]), debye_waller=False)
    pos = np.array([[0.0, 0.0], [3.3, -2.1], [-5.25, 7.7]])
    _run_sim(sim, 2.0, pos)
    t_ref, _ = oracle_slices(sim, 2.0)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 2.0, 1 / 3))
    assert sim.probes.shape == (len(pos), Ny, Nx)
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(len(pos))]
    _say(f"  S={sim.num_slices} CBED relL2 per probe {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-4


def test_exit_wave_and_incident_parity(gpu):
    _say("exit wave / incident 256^2")
    rec = random_block(7, 30.0, 16.0, 800)
    sim = MSAGPU(100, 0.02, rec, sampling=np.array([256, 256]), debye_waller=False)
    pos = np.array([[1.0, -1.0], [-4.5, 2.25], [0.0, 8.0]])
    _run_sim(sim, 2.0, pos, output="exit_wave")
    t_ref, _ = oracle_slices(sim, 2.0)
    psi_k, _ = oracle_probe(sim)
    psi0 = O.probe_stack(psi_k, pos.astype(np.float32), np.float32(sim.kmax))
    _, norm = O.probe_real(psi_k)
    ex = O.multislice(psi0, t_ref, oracle_pm(sim, 2.0, 1 / 3)) / np.sqrt(norm)
    assert rel_l2(sim.probes, ex) < 1e-4
    sim.multislice(transmit=False)
    assert rel_l2(sim.probes, psi0 / np.sqrt(norm)) < 2e-6


def test_fused_detectors_and_integrate(gpu):
    _say("detectors 256^2")
    rec = random_block(8, 32.0, 10.0, 600)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    sim.build_probe(sim_probe())
    sim.generate_probe_positions(grid_steps=np.array([3, 2]))
    sim.plan_simulation()
    dets = np.stack([sim.detector_mask(0.0, 25e-3), sim.detector_mask(40e-3, 0.9 * sim.max_ang),
                     sim.detector_mask(10e-3, 60e-3, mode="physical")])
    sim.multislice(detectors=dets)
    cb = sim.probes.astype(np.float64)
    ref = np.einsum("pyx,dyx->pd", cb, dets.astype(np.float64))
    np.testing.assert_allclose(sim.detector_signals, ref, rtol=1e-6)
    sig, det = sim.integrate_cbed(detector_params={"inner_angle": 0.0, "outer_angle": 25e-3})
    np.testing.assert_allclose(sig.ravel(), ref[:, 0], rtol=1e-6)
    assert sig.shape == (3, 2)


def test_unitarity_full_mask_config2_size(gpu):
    """Size-independent property at full size: with the band limit off, every
    slice is unitary, so sum(CBED) = 1 for each of B probes after 150 slices."""
    _say("unitarity 512^2 S=150")
    rec = structures.random_amorphous_crystal(seed=2)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([512, 512]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    assert sim.num_slices == 150
    sim.build_probe(sim_probe())
    sim.generate_probe_positions(grid_steps=np.array([8, 4]))
    sim.plan_simulation()
    sim.multislice(bandwidth=2.0)
    tot = sim.check_simulation()
    _say(f"  sum CBED after 150 slices: {np.min(tot):.7f} .. {np.max(tot):.7f}")
    np.testing.assert_allclose(tot, 1.0, rtol=0, atol=5e-5)


def test_config2_full_size_parity_two_probes(gpu):
    """Full config-2 workload (512^2, 150 slices, 73,728 atoms) for two raster
    points against the FP64 oracle."""
    _say("config2 full-size parity")
    rec = structures.random_amorphous_crystal(seed=2)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([512, 512]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    sim.build_probe(sim_probe())
    sim.generate_probe_positions(grid_steps=np.array([64, 64]))
    pos = sim.probe_positions[[0, 2080]]
    sim.probe_positions = pos
    sim.num_probes = np.int32(2)
    sim.grid_steps = np.array([2, 1])
    sim.plan_simulation()
    sim.multislice()
    t_ref, _ = oracle_slices(sim, 2.0)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 2.0, 1 / 3))
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(2)]
    _say(f"  config2 CBED relL2 {errs}")
    assert max(errs) < 1e-4


@pytest.mark.parametrize("N,S,bound", [(256, 300, 2e-5), (384, 200, 5e-5), (384, 300, 4.5e-5),
                                       (512, 300, 3.5e-5), (512, 1000, 6e-5), (768, 200, 5e-5),
                                       (768, 300, 4.5e-5), (1024, 300, 3.5e-5), (2048, 1000, 6e-5)])
def test_vacuum_round_off_growth(gpu, N, S, bound):
    """Size-independent property over thickness: in vacuum (t = 1) with the
    band limit off every slice is exactly unitary and diagonal in k, so the
    CBED after S slices equals the incident one; the difference is the
    accumulated round-off of the schedules the driver runs.  It grows
    linearly.  Up to 300 slices (200 for the radix-3 lengths 384, 768) the
    default schedules run, compact twiddles in some passes (at most about
    9e-8 per slice, 2e-7 for radix 3, when both passes are compact); beyond,
    with every twiddle held, about 4.5e-8 per slice for the 2-3 stage
    power-of-two transforms (1000 slices stay well inside the 1e-4 budget)
    and 1.2e-7 for the 4-5 stage radix-3 ones."""
    import torch

    from namsa_amd._lib import call, ptr, stream_handle

    _say(f"vacuum round-off N={N} S={S}")
    kmax, lam = 12.0, 0.0197
    st = stream_handle()
    psi_k = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ap = torch.empty((N, N), device="cuda")
    call("namsa_probe_k", N, N, np.float32(kmax), np.float32(0.030 / lam), np.float32(lam), np.float32(0),
         np.float32(0), np.float32(0), 1, np.float32(80.0), ptr(psi_k), ptr(ap), st)
    pm = torch.empty_like(psi_k)
    mask = torch.empty((N, N), device="cuda")
    call("namsa_propagator", N, N, np.float32(kmax), np.float32(10 * kmax), np.float32(lam), np.float32(1.0),
         ptr(pm), ptr(mask), st)
    t = torch.ones((S, N, N), dtype=torch.complex64, device="cuda")
    B = 2
    pos = torch.tensor([[0.0, 0.0], [3.3, -7.1]], dtype=torch.float32, device="cuda")
    ws = torch.empty((B, N, N), dtype=torch.complex64, device="cuda")
    cb = torch.empty((2, B, N, N), device="cuda")
    for i, s in enumerate((0, S)):
        call("namsa_multislice", ptr(psi_k), ptr(pos), B, N, N, np.float32(kmax), ptr(t), s, ptr(pm), 0, N, B,
             L.OUT_CBED, np.float32(1.0), ptr(ws), ptr(cb[i]), None, 0, None, None, st)
    c = cb.double().cpu().numpy()
    err = max(rel_l2(c[1, b], c[0, b]) for b in range(B))
    _say(f"  CBED change after {S} vacuum slices: {err:.2e}")
    assert err < bound


@pytest.mark.parametrize("N", [384, 512])
def test_thick_runs_use_the_exact_partner_schedules(gpu, N):
    """The fused driver's schedule switch, pinned bit for bit: at the
    threshold and one slice beyond it, a run with the default selection
    equals a run with the schedules autotune.driver_schedules names for that
    thickness selected explicitly."""
    import torch

    from namsa_amd import autotune
    from namsa_amd._lib import call, load, ptr, stream_handle

    lib = load()
    thr = lib.namsa_precise_slices(N)
    _say(f"thick-run schedule switch N={N} at {thr} and {thr + 1} slices")
    rng = np.random.default_rng(N)
    S = thr + 1
    st = stream_handle()
    c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).cuda()
    psi_k = c64((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N)
    t = c64(np.exp(1j * rng.uniform(0, 0.05, (S, N, N))))
    pm = c64(np.exp(1j * rng.uniform(0, 2 * np.pi, (N, N))))
    pos = torch.tensor([[0.0, 0.0], [2.5, -4.0]], dtype=torch.float32, device="cuda")
    B = 2
    ws = torch.empty((B, N, N), dtype=torch.complex64, device="cuda")

    def run(s):
        cb = torch.empty((B, N, N), device="cuda")
        call("namsa_multislice", ptr(psi_k), ptr(pos), B, N, N, np.float32(6.0), ptr(t), s, ptr(pm), 0, N, B,
             L.OUT_CBED, np.float32(1.0), ptr(ws), ptr(cb), None, 0, None, None, st)
        return cb.cpu().numpy()

    keep = (autotune.selected(N, 0), autotune.selected(N, 1))
    try:
        for s in (thr, thr + 1):
            default = run(s)
            rows, cols = autotune.driver_schedules(N, s)
            autotune.select(N, 0, rows)
            autotune.select(N, 1, cols)
            explicit = run(s)
            autotune.select(N, 0, keep[0])
            autotune.select(N, 1, keep[1])
            _say(f"  S={s}: {autotune.schedules(N, 0)[rows]} / {autotune.schedules(N, 1)[cols]}")
            assert np.array_equal(default, explicit)
        assert autotune.driver_schedules(N, thr) == keep
    finally:
        autotune.select(N, 0, keep[0])
        autotune.select(N, 1, keep[1])


def test_polycrystal_2048_parity(gpu):
    """A thin slab of the seeded 2048^2 polycrystal (grain boundaries, four
    elements, 300 kV, 30 mrad) against the oracle, one probe near a triple
    junction region and one in a grain."""
    from namsa_amd import structures

    _say("polycrystal 2048^2, 6 slices")
    rec, info = structures.polycrystal(seed=5, thickness=6.0)
    sim = MSAGPU(300, 0.030, rec, sampling=np.array([2048, 2048]), debye_waller=False)
    pos = np.array([[0.0, 0.0], [-20.5, 31.25]])
    _run_sim(sim, 1.0, pos)
    t_ref, _ = oracle_slices(sim, 1.0)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 1.0, 1 / 3))
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(len(pos))]
    _say(f"  atoms {info['atoms']} S={sim.num_slices} CBED relL2 {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-4


def test_vacuum_closed_form_2048(gpu):
    """No slices (S=0) and transparent slices both give mask-free / masked |psi_k|^2."""
    import torch

    _say("vacuum 2048^2")
    N = 2048
    rec = random_block(4, 128.0, 4.0, 50)
    sim = MSAGPU(300, 0.03, rec, sampling=np.array([N, N]), debye_waller=False)
    sim.setup_device(0)
    sim.slice_t = 1.0
    sim.build_probe(sim_probe())
    psi_k, _ = oracle_probe(sim)
    pos = torch.tensor([[0.0, 0.0], [10.0, -20.0]], dtype=torch.float32, device=gpu)
    cb = torch.empty((2, N, N), dtype=torch.float32, device=gpu)
    ws = torch.empty((2, N, N), dtype=torch.complex64, device=gpu)
    st = L.stream_handle()
    L.call("namsa_multislice", L.ptr(sim.psi_k_device), L.ptr(pos), 2, N, N, np.float32(sim.kmax), None, 0,
           None, 0, 0, 0, L.OUT_CBED, np.float32(1 / sim.normalization), L.ptr(ws), L.ptr(cb), None, 0, None,
           None, st)
    expect = np.abs(psi_k) ** 2 / np.sum(np.abs(psi_k) ** 2)
    for i in range(2):
        assert rel_l2(cb[i].cpu().numpy(), expect) < 1e-5
    ones = torch.ones((3, N, N), dtype=torch.complex64, device=gpu)
    pm = sim.build_propagator(1 / 3)
    L.call("namsa_multislice", L.ptr(sim.psi_k_device), L.ptr(pos), 2, N, N, np.float32(sim.kmax), L.ptr(ones), 3,
           L.ptr(pm), sim.band[0], sim.band[1], 1, L.OUT_CBED, np.float32(1 / sim.normalization), L.ptr(ws),
           L.ptr(cb), None, 0, None, None, st)
    masked = expect * (np.abs(oracle_pm(sim, 1.0, 1 / 3)) > 0)
    for i in range(2):
        assert rel_l2(cb[i].cpu().numpy(), masked) < 1e-5


def test_config1_si111_parity(gpu):
    """Config-1 workload: Si[111] 256^2, ~104 slices of 1.92 A, one centred probe, 100 kV."""
    _say("config1 Si[111]")
    rec = structures.si111_supercell()
    sim = MSAGPU(100, 0.020, rec, sampling=np.array([256, 256]), debye_waller=False)
    _run_sim(sim, 1.92, np.array([[0.0, 0.0]]))
    assert sim.num_slices == 104
    t_ref, _ = oracle_slices(sim, 1.92)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, np.zeros((1, 2), np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 1.92, 1 / 3))
    err = rel_l2(sim.probes[0], cb_ref[0])
    _say(f"  config1 CBED relL2 {err:.2e}")
    assert err < 1e-4
