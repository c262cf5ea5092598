"""Host-side logic of the drop-in class against the oracle (CPU only, no GPU calls)."""
import numpy as np
import pytest

from namsa_amd import MSAGPU, structures
from namsa_amd.optics import sigma_int, voltage2Lambda
from oracle import msa_oracle as O


def _sim(rec, N=256, kv=200.0, semi=0.025):
    return MSAGPU(kv, semi, rec, sampling=np.array([N, N]), debye_waller=False)


@pytest.fixture(scope="module")
def config2():
    return structures.random_amorphous_crystal(seed=2)


def test_optics_match_oracle():
    for v in (60e3, 100e3, 200e3, 300e3):
        assert voltage2Lambda(v) == O.voltage2lambda(v)
        assert sigma_int(v) == pytest.approx(O.sigma_int(v), rel=1e-15)


def test_geometry_matches_oracle(config2):
    sim = _sim(config2, N=512)
    g = O.geometry(sim.supercell_xyz, (512, 512), 200.0)
    np.testing.assert_array_equal(sim.dims, g["dims"])
    assert sim.kmax == g["kmax"] and sim.max_ang == g["max_ang"]
    np.testing.assert_array_equal(sim.pix_size, g["pix_size"])
    assert sim.sigma == pytest.approx(g["sigma"], rel=1e-15)


def test_config2_structure_invariants(config2):
    from scipy.spatial import cKDTree

    xyz = np.stack([config2["x"], config2["y"], config2["z"]], 1).astype(np.float64)
    assert len(config2) == 73728                       # 64*64*300*0.06
    assert set(np.unique(config2["atomic_number"])) == {6, 14, 29, 79}
    d, _ = cKDTree(xyz).query(xyz, k=2)
    assert d[:, 1].min() >= 1.2 - 1e-4
    assert xyz[:, 2].max() - xyz[:, 2].min() == pytest.approx(300.0, abs=1e-4)
    assert int(np.floor(np.float64(np.float32(xyz[:, 2].max() - xyz[:, 2].min())) / 2.0)) == 150


def test_si111_structure():
    rec = structures.si111_supercell()
    xyz = np.stack([rec["x"], rec["y"], rec["z"]], 1)
    assert set(rec["atomic_number"]) == {14}
    dims = xyz.max(0) - xyz.min(0)
    assert 36 < dims[0] < 39 and 37 < dims[1] < 40 and 199 < dims[2] < 200
    assert int(np.floor(np.float64(dims[2]) / 1.92)) == 104
    from scipy.spatial import cKDTree

    d, _ = cKDTree(xyz.astype(np.float64)).query(xyz.astype(np.float64), k=2)
    assert d[:, 1].min() == pytest.approx(5.431 * np.sqrt(3) / 4, rel=1e-4)  # Si-Si bond


def test_polycrystal_structure():
    """Seeded Voronoi polycrystal: deterministic, fills the box, grain lattices
    keep their nearest-neighbour distances, boundaries keep >= dmin."""
    from scipy.spatial import cKDTree

    rec, info = structures.polycrystal(seed=5, lateral=40.0, thickness=30.0, slab=10.0)
    rec2, _ = structures.polycrystal(seed=5, lateral=40.0, thickness=30.0, slab=10.0)
    np.testing.assert_array_equal(rec, rec2)
    xyz = np.stack([rec["x"], rec["y"], rec["z"]], 1).astype(np.float64)
    assert info["grains"] == ["Si", "GaAs", "Au", "Si"]
    assert set(np.unique(rec["atomic_number"])) <= {14, 31, 33, 79}
    assert xyz.min() >= 0 and xyz[:, :2].max() < 40.0 and xyz[:, 2].max() == pytest.approx(30.0)
    d, _ = cKDTree(xyz).query(xyz, k=2)
    assert d[:, 1].min() >= 1.2 - 1e-4
    # density between Au-fcc (0.059) and GaAs (0.044) per cubic Angstrom
    assert 0.04 < len(rec) / (40 * 40 * 30) < 0.065
    # inside Au grains the fcc nearest-neighbour distance a/sqrt(2) dominates
    au = xyz[rec["atomic_number"] == 79]
    da, _ = cKDTree(au).query(au, k=2)
    assert np.median(da[:, 1]) == pytest.approx(4.078 / np.sqrt(2), rel=1e-4)


def test_site_table_matches_oracle_binning(config2):
    sim = _sim(config2, N=512)
    sim.unique_Z = np.unique(sim.supercell_Z)
    sites, offs = sim.site_table(2.0)
    ref, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, 2.0, sim.pix_size, sim.dims)
    assert len(offs) - 1 == len(ref) == 150
    for s in range(len(ref)):
        np.testing.assert_array_equal(sites[offs[s]:offs[s + 1]], ref[s])


def test_site_table_boundaries_and_negative_coordinates():
    xyz = np.array([[-0.3, 1.0, 0.0], [2.0, -0.9, 1.0], [5.0, 5.0, 2.0], [9.9, 9.9, 4.0], [0.0, 0.0, 8.0],
                    [10.0, 10.0, 3.999999]], np.float32)
    rec = structures.make_records(xyz, [6, 14, 6, 79, 14, 6])
    sim = MSAGPU(100.0, 0.02, rec, sampling=np.array([128, 128]), debye_waller=False)
    sim.unique_Z = np.unique(sim.supercell_Z)
    sites, offs = sim.site_table(2.0)
    ref, _ = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, 2.0, sim.pix_size, sim.dims)
    for s in range(len(ref)):
        np.testing.assert_array_equal(sites[offs[s]:offs[s + 1]], ref[s])


def test_probe_positions_match_oracle(config2):
    sim = _sim(config2, N=512)
    sim.generate_probe_positions(grid_steps=np.array([64, 64]))
    ref, gs = O.probe_positions(sim.dims, grid_steps=np.array([64, 64]))
    np.testing.assert_array_equal(sim.probe_positions, ref)
    assert sim.num_probes == 4096 and list(sim.grid_steps) == [64, 64]
    sim.generate_probe_positions(probe_step=np.array([2.0, 4.0]))
    ref2, gs2 = O.probe_positions(sim.dims, probe_step=np.array([2.0, 4.0]))
    np.testing.assert_array_equal(sim.probe_positions, ref2)


def test_detector_masks(config2):
    sim = _sim(config2, N=512)
    ref = O.detector_mask_reference((512, 512), sim.kmax, sim.Lambda, 40e-3, 200e-3)
    np.testing.assert_array_equal(sim.detector_mask(40e-3, 200e-3), ref)
    phys = sim.detector_mask(0.0, 25e-3, mode="physical")
    theta = O.k_radius(512, 512, np.float32(sim.kmax)) * sim.Lambda
    np.testing.assert_array_equal(phys, (theta < 25e-3).astype(np.float32))
    with pytest.raises(ValueError):
        sim.detector_mask(40e-3, 0.3)


def test_table_shape_rule(config2):
    sim = _sim(config2, N=512)
    shape = tuple(int(np.floor(8 / p)) for p in sim.pix_size)
    assert shape == O.table_shape(sim.pix_size) == (64, 64)


def test_product_does_not_import_oracle():
    import pathlib

    for f in pathlib.Path("namsa_amd").rglob("*.py"):
        src = f.read_text()
        assert "import oracle" not in src and "from oracle" not in src, f


def test_probe_group_policy(monkeypatch, tmp_path):
    # waves of one group fill the measured share of the 256 MB last-level
    # cache (no tuner cache present); grids without a measurement use 3/4
    from namsa_amd import msa

    monkeypatch.setenv("NAMSA_TUNE_CACHE", str(tmp_path / "none.json"))
    llc = 256 << 20
    for n, share in msa._GROUP_SHARE.items():
        assert msa.cache_group(n, n) == int(llc * share // (8 * n * n))
    assert msa.cache_group(512, 512) == 160 and msa.cache_group(2048, 2048) == 6
    assert msa.cache_group(256, 512) == int(llc * 0.75 // (8 * 256 * 512))
    assert msa.cache_group(8192, 8192) == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py reports one MI355X; --gpus above 1 exits with a message, before any GPU work."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    res = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120, cwd=root)
    assert res.returncode != 0
    assert "one MI355X" in (res.stdout + res.stderr) and res.stdout.strip() == ""


class _FakeSlices:
    """Stands in for the device slices: plan_simulation only reads numel()."""

    def __init__(self, S, N):
        self.n = S * N * N

    def numel(self):
        return self.n


@pytest.mark.parametrize("N,P,S,free_gib,expect_max", [
    (512, 4096, 150, 280.0, 4096),      # config 2 fits at once: the output stays on the device
    (2048, 4096, 1000, 280.0, 1728),    # config 5: 4096 -> 3072 -> 2304 -> 1728 (x3/4 each), 149.7 of 150.3 GB
    (512, 4096, 150, 4.0, 306),         # a small card (4 GiB free): phases of 306
])
def test_memory_plan_and_phases_with_faked_free_memory(monkeypatch, tmp_path, N, P, S, free_gib, expect_max):
    """plan_simulation sizes the phases from torch.cuda.mem_get_info as the
    reference sizes them from cuda.mem_get_info (msa.py:445-461): 16 N^2
    bytes per position, one probe group of state and the resident slices
    within half the free memory, shrinking by 3/4.  phase_plan then covers
    every position once, in order, never more than max_probes at a time."""
    import torch

    from namsa_amd import msa

    monkeypatch.setenv("NAMSA_TUNE_CACHE", str(tmp_path / "none.json"))
    free = int(free_gib * (1 << 30))
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (free, 288 << 30))
    xyz = np.array([[0.0, 0.0, 0.0], [64.0, 64.0, 300.0], [10.0, 20.0, 30.0]])
    sim = _sim(structures.make_records(xyz, np.array([14, 14, 79])), N=N)
    sim.device = torch.device("cpu")          # host-only: no device is touched
    sim._t = _FakeSlices(S, N)
    sim.num_probes = np.int32(P)
    G = msa.cache_group(N, N)
    got = sim.plan_simulation()
    assert got == sim.max_probes == expect_max
    plane = N * N
    need = lambda n: n * 16 * plane + min(n, G) * 8 * plane + S * plane * 8
    assert need(got) <= 0.5 * free
    assert got == P or need(int(got / 0.75) + 1) > 0.5 * free   # the shrink stopped at the first fit
    for sink, batch in ((False, None), (True, None), (False, 100)):
        chunk, resident = msa.phase_plan(P, got, G, sink, batch)
        assert 1 <= chunk <= got and (batch is None or chunk <= batch)
        assert resident == (not sink and P <= got)
        starts = list(range(0, P, chunk))
        order = np.concatenate([np.arange(s, min(P, s + chunk)) for s in starts])
        np.testing.assert_array_equal(order, np.arange(P))     # in order, each position once
        if sink and batch is None:
            assert chunk % G == 0 or chunk == got


def test_phase_plan_small_cases():
    from namsa_amd.msa import phase_plan

    assert phase_plan(36, 1000, 160, sink=True, batch_size=8) == (8, False)
    assert phase_plan(36, 1000, 160) == (36, True)
    assert phase_plan(36, 10, 160) == (10, False)
    assert phase_plan(4096, 4096, 160, sink=True) == (480, False)   # 160 * (26 // 8): ~8 chunks
    assert phase_plan(1, 1, 6) == (1, True)
