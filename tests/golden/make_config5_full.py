"""Writes tests/golden/config5_full_samples.npz: the FP64 oracle's CBEDs of the
2048^2 polycrystal workload after its full 999 slices (seed-5 polycrystal,
dz = 1 A, 300 kV, 30 mrad, bandwidth 1/3) at the two probe positions of
tests/full_size_config5.py, sampled at 65536 seeded pixels of the band window
per position, with each CBED's total and squared norm.

The full CBEDs (2 x 2048^2) are too large to keep; a rel-L2 over a uniform
sample of the window estimates the full one (the CBED is zero outside it).
Run on a CPU (about an hour: the oracle streams the slices one at a time):

    python tests/golden/make_config5_full.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import msa_oracle as O  # noqa: E402
from tests.full_size_config5 import POS, _sim  # noqa: E402

N_SAMPLES = 65536
SEED = 5


def main(out=ROOT / "tests" / "golden" / "config5_full_samples.npz"):
    sim, info = _sim()
    Ny, Nx = 2048, 2048
    kirk = np.load(ROOT / "namsa_amd" / "data" / "kirkland_params.npy")
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, 1.0, sim.pix_size, sim.dims)
    tables = np.stack([O.atomic_table(kirk[z - 1], sim.pix_size) for z in uniq])
    psi_k, _ = O.probe_k(Ny, Nx, np.float32(sim.kmax), np.float32(sim.semi_ang / sim.Lambda),
                         np.float32(sim.Lambda), 0.0, 0.0, 0.0, smooth=True)
    pm = O.propagator_mask(Ny, Nx, np.float32(sim.kmax), np.float32(1 / 3 * sim.kmax), np.float32(sim.Lambda),
                           np.float32(1.0))
    psi = O.probe_stack(psi_k, POS.astype(np.float32), np.float32(sim.kmax))
    _, norm = O.probe_real(psi_k)
    t0 = time.time()
    for i, st in enumerate(sites):
        V = O.splat_slices(tables, [st], Ny, Nx)
        psi = O.multislice(psi, O.transmission(V, np.float32(sim.sigma)), pm)
        if i % 50 == 0:
            print(f"oracle slice {i}/{len(sites)} ({time.time() - t0:.0f} s)", flush=True)
    cb = O.readout(psi, norm)
    if len(sys.argv) > 1:   # optional path for the full patterns
        np.save(sys.argv[1], cb)
    # the band window: rows and columns where the propagator mask is open
    mask = np.abs(pm) > 0
    rows, cols = np.nonzero(mask.any(axis=1))[0], np.nonzero(mask.any(axis=0))[0]
    y0, y1, x0, x1 = int(rows[0]), int(rows[-1]) + 1, int(cols[0]), int(cols[-1]) + 1
    # outside the window the oracle's CBED is FP64 round-off of an exactly band-limited spectrum
    outside = cb.copy()
    outside[:, y0:y1, x0:x1] = 0
    assert np.abs(outside).max() < 1e-12 * np.abs(cb).max(), np.abs(outside).max()
    rng = np.random.default_rng(SEED)
    iy = rng.integers(y0, y1, N_SAMPLES).astype(np.int16)
    ix = rng.integers(x0, x1, N_SAMPLES).astype(np.int16)
    np.savez_compressed(out, pos=POS, slices=len(sites), window=np.array([y0, y1, x0, x1]), iy=iy, ix=ix,
                        samples=cb[:, iy, ix].astype(np.float64), total=cb.sum(axis=(1, 2)),
                        norm2=(cb.astype(np.float64) ** 2).sum(axis=(1, 2)), atoms=info["atoms"])
    print(f"wrote {out}: {len(sites)} slices, window {y0}:{y1} x {x0}:{x1}, totals {cb.sum(axis=(1, 2))}, "
          f"{time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
