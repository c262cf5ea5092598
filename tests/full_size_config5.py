"""Full-thickness parity of the 2048^2 polycrystal workload (999 slices).

Too long for the test suite (the FP64 oracle needs ~2000 FFTs of 2048^2), so
it runs in two steps:

  on the GPU:   python tests/full_size_config5.py gpu  c5_full.npz
  on any CPU:   python tests/full_size_config5.py oracle c5_full.npz [oracle.npz]
  later runs:   python tests/full_size_config5.py compare c5_full.npz oracle.npz

The first writes the GPU CBEDs of two probe positions (MSAGPU, dz = 1 A,
bandwidth 1/3); the second rebuilds the same structure, streams the oracle
through the slices one at a time (its full slice stack would need 67 GB) and
prints the CBED rel-L2 per position.  The 1e-4 bar applies.
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from namsa_amd import structures  # noqa: E402
from namsa_amd.msa import DEFAULT_PROBE, MSAGPU  # noqa: E402

POS = np.array([[0.0, 0.0], [-20.5, 31.25]])   # (y, x) Angstrom from the centre


def _sim():
    rec, info = structures.polycrystal(seed=5)
    return MSAGPU(300, 0.030, rec, sampling=np.array([2048, 2048]), debye_waller=False), info


def run_gpu(out):
    sim, info = _sim()
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(1.0)
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.probe_positions = POS.copy()
    sim.num_probes = np.int32(len(POS))
    sim.grid_steps = np.array([len(POS), 1])
    sim.plan_simulation()
    t0 = time.time()
    sim.multislice(bandwidth=1 / 3)
    cb = sim.probes
    print(f"gpu: {info['atoms']} atoms, {int(sim.num_slices)} slices, {time.time() - t0:.1f} s, "
          f"sum CBED {cb.sum(axis=(1, 2))}", flush=True)
    np.savez_compressed(out, cbed=cb, pos=POS, slices=int(sim.num_slices))


def compare(cb_gpu, cb):
    for b in range(len(POS)):
        e = np.linalg.norm(cb_gpu[b] - cb[b]) / np.linalg.norm(cb[b])
        print(f"position {POS[b]}: CBED rel-L2 {e:.3e}  (sum GPU {cb_gpu[b].sum():.6f}, oracle {cb[b].sum():.6f})")


def run_compare(path, oracle_path):
    compare(np.load(path)["cbed"], np.load(oracle_path)["cbed"])


def run_oracle(path, save=None):
    from oracle import msa_oracle as O

    data = np.load(path)
    cb_gpu = data["cbed"]
    sim, info = _sim()
    Ny, Nx = 2048, 2048
    kirk = np.load(ROOT / "namsa_amd" / "data" / "kirkland_params.npy")
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, 1.0, sim.pix_size, sim.dims)
    assert len(sites) == int(data["slices"])
    tables = np.stack([O.atomic_table(kirk[z - 1], sim.pix_size) for z in uniq])
    psi_k, _ = O.probe_k(Ny, Nx, np.float32(sim.kmax), np.float32(sim.semi_ang / sim.Lambda),
                      np.float32(sim.Lambda), 0.0, 0.0, 0.0, smooth=True)
    # the default probe applies the Scherzer defocus for C3 = 0, i.e. C1 = 0
    pm = O.propagator_mask(Ny, Nx, np.float32(sim.kmax), np.float32(1 / 3 * sim.kmax), np.float32(sim.Lambda),
                           np.float32(1.0))
    psi = O.probe_stack(psi_k, POS.astype(np.float32), np.float32(sim.kmax))
    _, norm = O.probe_real(psi_k)
    t0 = time.time()
    for i, st in enumerate(sites):
        V = O.splat_slices(tables, [st], Ny, Nx)
        psi = O.multislice(psi, O.transmission(V, np.float32(sim.sigma)), pm)
        if i % 100 == 0:
            print(f"oracle slice {i}/{len(sites)} ({time.time() - t0:.0f} s)", flush=True)
    cb = O.readout(psi, norm)
    if save:
        np.savez_compressed(save, cbed=cb, pos=POS)
    compare(cb_gpu, cb)


if __name__ == "__main__":
    {"gpu": run_gpu, "oracle": run_oracle, "compare": run_compare}[sys.argv[1]](*sys.argv[2:])
