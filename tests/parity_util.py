"""Shared helpers for the GPU parity tests: run the FP64 oracle on exactly the
inputs a drop-in simulation object uses."""
import numpy as np

from oracle import msa_oracle as O

KIRK = np.load("namsa_amd/data/kirkland_params.npy")


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def oracle_slices(sim, dz):
    """FP64 transmission slices for the simulation's atoms (independent tables)."""
    Ny, Nx = int(sim.sampling[0]), int(sim.sampling[1])
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, dz, sim.pix_size, sim.dims)
    tables = np.stack([O.atomic_table(KIRK[z - 1], sim.pix_size) for z in uniq])
    V = O.splat_slices(tables, sites, Ny, Nx)
    return O.transmission(V, np.float32(sim.sigma)), tables


def oracle_probe(sim, smooth=True, C=(0.0, 0.0, 0.0)):
    Ny, Nx = int(sim.sampling[0]), int(sim.sampling[1])
    return O.probe_k(Ny, Nx, np.float32(sim.kmax), np.float32(sim.semi_ang / sim.Lambda),
                     np.float32(sim.Lambda), *C, smooth=smooth)


def oracle_pm(sim, dz, bandwidth):
    Ny, Nx = int(sim.sampling[0]), int(sim.sampling[1])
    return O.propagator_mask(Ny, Nx, np.float32(sim.kmax), np.float32(bandwidth * sim.kmax),
                             np.float32(sim.Lambda), np.float32(dz))


def random_block(seed, lateral, thickness, n_atoms, zset=(6, 14, 29, 79)):
    from namsa_amd import structures

    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n_atoms, 3)) * np.array([lateral, lateral, thickness])
    xyz[0] = 0.0
    xyz[1] = [lateral, lateral, thickness]
    return structures.make_records(xyz, rng.choice(np.asarray(zset), n_atoms))
