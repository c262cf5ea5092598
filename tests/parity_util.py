"""Shared helpers for the GPU parity tests: run the FP64 oracle on exactly the
inputs a drop-in simulation object uses."""
import numpy as np

from oracle import msa_oracle as O

KIRK = np.load("namsa_amd/data/kirkland_params.npy")


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def oracle_geometry(sim):
    """The oracle's geometry for the simulation's atoms, sampling and energy,
    derived without the product's fields; the product's dims, kmax, pix_size
    (py, px), max_ang and sigma must equal it."""
    g = O.geometry(sim.supercell_xyz, sim.sampling, sim.E)
    np.testing.assert_array_equal(sim.dims, g["dims"])
    np.testing.assert_array_equal(sim.pix_size, g["pix_size"])
    assert sim.kmax == g["kmax"] and sim.max_ang == g["max_ang"]
    assert abs(sim.sigma / g["sigma"] - 1) < 1e-15 and sim.Lambda == g["lam"]
    return g


def oracle_slices(sim, dz):
    """FP64 transmission slices for the simulation's atoms (independent tables)."""
    g = oracle_geometry(sim)
    Ny, Nx = (int(v) for v in g["sampling"])
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, dz, g["pix_size"], g["dims"])
    tables = np.stack([O.atomic_table(KIRK[z - 1], g["pix_size"]) for z in uniq])
    V = O.splat_slices(tables, sites, Ny, Nx)
    return O.transmission(V, np.float32(g["sigma"])), tables


def oracle_probe(sim, smooth=True, C=(0.0, 0.0, 0.0)):
    g = oracle_geometry(sim)
    Ny, Nx = (int(v) for v in g["sampling"])
    return O.probe_k(Ny, Nx, np.float32(g["kmax"]), np.float32(sim.semi_ang / g["lam"]),
                     np.float32(g["lam"]), *C, smooth=smooth)


def oracle_pm(sim, dz, bandwidth):
    g = oracle_geometry(sim)
    Ny, Nx = (int(v) for v in g["sampling"])
    return O.propagator_mask(Ny, Nx, np.float32(g["kmax"]), np.float32(bandwidth * g["kmax"]),
                             np.float32(g["lam"]), np.float32(dz))


def random_block(seed, lateral, thickness, n_atoms, zset=(6, 14, 29, 79)):
    from namsa_amd import structures

    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n_atoms, 3)) * np.array([lateral, lateral, thickness])
    xyz[0] = 0.0
    xyz[1] = [lateral, lateral, thickness]
    return structures.make_records(xyz, rng.choice(np.asarray(zset), n_atoms))
