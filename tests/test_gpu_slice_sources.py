"""Alternate slice sources on MI355X against the FP64 oracle: slices
integrated from a sampled 3D potential (MSA/MSAGPU.integrate_potential_slices,
msa.py:373-424, :568-578) and the thin-object approximation
(utils.py:130-135)."""
import sys

import numpy as np
import pytest

import namsa_amd._lib as L
from namsa_amd import DEFAULT_PROBE, MSAGPU
from oracle import msa_oracle as O
from tests.parity_util import random_block, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gaussian_potential(nz, ny, nx, dz, dy, dx, seed, n_atoms=40):
    """A smooth synthetic 'ab initio' potential: Gaussian blobs on the grid."""
    rng = np.random.default_rng(seed)
    z, y, x = np.arange(nz) * dz, np.arange(ny) * dy, np.arange(nx) * dx
    pot = np.zeros((nz, ny, nx))
    for _ in range(n_atoms):
        cz, cy, cx = rng.uniform(0, z[-1]), rng.uniform(0, y[-1]), rng.uniform(0, x[-1])
        a, w = rng.uniform(5, 40), rng.uniform(0.3, 0.8)
        pot -= a * np.exp(-((z - cz)[:, None, None] ** 2 + (y - cy)[None, :, None] ** 2
                            + (x - cx)[None, None, :] ** 2) / (2 * w * w))
    return pot, (z, y, x)


def test_integrate_slices_kernel_matches_oracle(gpu):
    import torch

    _say("namsa_integrate_slices vs oracle, 256^2, 4 slices of 9 samples")
    pot, grid = _gaussian_potential(37, 256, 256, 0.25, 0.1, 0.1, seed=4)
    V_ref, info = O.integrate_slices(pot, grid, 2.0)
    S, d = info["num_slices"], info["d"]
    assert (S, d) == (4, 9)
    pot_d = torch.from_numpy(np.ascontiguousarray(pot[: S * d])).to(gpu)
    z_d = torch.from_numpy(np.ascontiguousarray(grid[0][: S * d])).to(gpu)
    t = torch.empty((S, 256, 256), dtype=torch.complex64, device=gpu)
    v = torch.empty((S, 256, 256), dtype=torch.float64, device=gpu)
    sigma = 7.2884e-4
    L.call("namsa_integrate_slices", L.ptr(pot_d), L.ptr(z_d), S, d, 256, 256, sigma, L.ptr(t), L.ptr(v),
           L.stream_handle())
    V = v.cpu().numpy()
    # same FP64 terms in the same order, no contraction: equal to rounding of the sum
    np.testing.assert_allclose(V, V_ref, rtol=0, atol=1e-12 * np.abs(V_ref).max())
    t_ref = np.exp(1j * sigma * V_ref).astype(np.complex64)
    assert np.abs(t.cpu().numpy() - t_ref).max() < 2.5e-7      # one float32 ulp of a unit phasor


def test_integrate_potential_slices_drop_in(gpu):
    """The drop-in method on a non-square plane with a z crop: simulation
    parameters, returned V and the transmission slices against the oracle,
    then the CBED of a multislice through those slices (1e-4)."""
    _say("MSAGPU.integrate_potential_slices: 256 x 260 plane, 41 z samples")
    pot, grid = _gaussian_potential(41, 260, 256, 0.2, 0.1, 0.1, seed=5)
    sim = MSAGPU(200, 0.025, random_block(5, 20.0, 8.0, 50), sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    V = sim.integrate_potential_slices(pot, grid, 2.0, output=True)
    V_ref, info = O.integrate_slices(pot, grid, 2.0)
    pr = O.integration_params(V_ref.shape, grid, sim.Lambda)
    assert int(sim.num_slices) == info["num_slices"] == 4 and info["z_used"] == 40
    assert list(sim.sampling) == list(pr["sampling"]) == [256, 256]
    np.testing.assert_allclose(sim.dims, pr["dims"])
    assert sim.kmax == pytest.approx(pr["kmax"], rel=1e-15)
    np.testing.assert_allclose(sim.pix_size, pr["pix_size"])
    np.testing.assert_allclose(V, V_ref, rtol=0, atol=1e-12 * np.abs(V_ref).max())
    t_ref = np.exp(1j * sim.sigma * V_ref).astype(np.complex64)
    assert np.abs(sim.potential_slices - t_ref).max() < 2.5e-7

    pos = np.array([[0.0, 0.0], [2.5, -3.75]])
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.probe_positions = pos.copy()
    sim.num_probes = np.int32(len(pos))
    sim.grid_steps = np.array([len(pos), 1])
    sim.plan_simulation()
    sim.multislice(bandwidth=1 / 3)
    psi_k, _ = O.probe_k(256, 256, np.float32(sim.kmax), np.float32(sim.semi_ang / sim.Lambda),
                         np.float32(sim.Lambda))
    pm = O.propagator_mask(256, 256, np.float32(sim.kmax), np.float32(sim.kmax / 3), np.float32(sim.Lambda),
                           np.float32(2.0))
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), np.exp(1j * sim.sigma * V_ref), pm)
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(len(pos))]
    _say(f"  CBED relL2 {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-4


def test_thin_object_matches_oracle(gpu):
    from namsa_amd.utils import thin_object_approx

    _say("thin_object_approx vs oracle, 512^2, 3 slices, 2 probes")
    rng = np.random.default_rng(6)
    N = 512
    V = rng.uniform(-0.4, 0.4, (3, N, N))
    probes = ((rng.normal(size=(2, N, N)) + 1j * rng.normal(size=(2, N, N))) / N).astype(np.complex64)
    got = thin_object_approx(V, probes)
    ref = O.thin_object(V, probes)
    assert got.shape == (2, N, N) and got.dtype == np.float32
    err = rel_l2(got, ref)
    _say(f"  relL2 {err:.2e}")
    assert err < 3e-6        # FP32 rounding of an N^2-point transform
    one = thin_object_approx(V, probes[1])
    assert one.shape == (N, N) and rel_l2(one, ref[1]) < 3e-6


@pytest.mark.parametrize("material", ["Si", "GaN"])
def test_cif_supercell_multislice(gpu, material, tmp_path):
    """Structure front-end feeding the GPU path: Si.cif (192 operations, a
    22 x 22 x 16 Å [110]/[001] box) and hexagonal GaN (P6_3mc, 20 x 20 x 12 Å)
    -> unit cell -> supercell -> MSAGPU slices and CBEDs, against the oracle's
    slices and CBEDs for the same records."""
    import os

    from namsa_amd.supercell import SupercellBuilder
    from tests.parity_util import oracle_pm, oracle_probe, oracle_slices
    from tests.test_supercell import GAN_CIF

    _say(f"{material} CIF -> SupercellBuilder -> MSAGPU, 256^2")
    if material == "Si":
        sp = SupercellBuilder(os.path.join(os.path.dirname(__file__), "golden", "cif_files", "Si.cif"))
        sp.build_unit_cell()
        sp.make_orthogonal_supercell(supercell_size=np.array([22.0, 22.0, 16.0]), projec_1=np.array([1, 1, 0]),
                                     projec_2=np.array([0, 0, 1]))
    else:
        path = tmp_path / "gan.cif"
        path.write_text(GAN_CIF)
        sp = SupercellBuilder(str(path))
        sp.build_unit_cell()
        sp.make_orthogonal_supercell(supercell_size=np.array([20.0, 20.0, 12.0]))
    sim = MSAGPU(200, 0.025, sp.supercell_sites, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    pos = np.array([[0.0, 0.0], [1.92, -0.96]])
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.probe_positions = pos.copy()
    sim.num_probes = np.int32(len(pos))
    sim.grid_steps = np.array([len(pos), 1])
    sim.plan_simulation()
    sim.multislice(bandwidth=1 / 3)
    t_ref, _ = oracle_slices(sim, 2.0)
    psi_k, _ = oracle_probe(sim)
    cb_ref = O.cbed(psi_k, pos.astype(np.float32), np.float32(sim.kmax), t_ref, oracle_pm(sim, 2.0, 1 / 3))
    errs = [rel_l2(sim.probes[i], cb_ref[i]) for i in range(len(pos))]
    _say(f"  {len(sp.supercell_sites)} atoms, S={sim.num_slices}, CBED relL2 {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-4
