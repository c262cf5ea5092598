"""Full-thickness parity of the 2048^2 polycrystal workload (config 5, 999
slices of dz = 1 A, 300 kV, 30 mrad, bandwidth 1/3) through the drop-in class,
against the FP64 oracle's CBEDs of the same two probe positions
(tests/golden/config5_full_samples.npz, written by
tests/golden/make_config5_full.py).

The oracle CBEDs are kept as 65536 seeded pixel samples of the band window per
position (the CBED is zero outside it) plus each pattern's total; the CBED
rel-L2 over the samples estimates the full one.  Bar: 1e-4 (measured
3.8e-5 on the full patterns).  The run takes the thick-run schedules (999 >
300 slices: every twiddle held)."""
from pathlib import Path

import numpy as np
import pytest

from tests.full_size_config5 import POS, _sim

pytestmark = pytest.mark.gpu
FIXTURE = Path(__file__).resolve().parent / "golden" / "config5_full_samples.npz"


def test_config5_full_thickness_matches_oracle(gpu):
    if not FIXTURE.exists():
        pytest.skip("tests/golden/config5_full_samples.npz not generated")
    ref = np.load(FIXTURE)
    sim, info = _sim()
    assert info["atoms"] == int(ref["atoms"])
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(1.0)
    assert int(sim.num_slices) == int(ref["slices"])
    from namsa_amd import DEFAULT_PROBE

    sim.build_probe(dict(DEFAULT_PROBE))
    sim.probe_positions = POS.copy()
    sim.num_probes = np.int32(len(POS))
    sim.grid_steps = np.array([len(POS), 1])
    sim.plan_simulation()
    sim.multislice(bandwidth=1 / 3)
    cb = sim.probes
    iy, ix = ref["iy"].astype(np.int64), ref["ix"].astype(np.int64)
    for b in range(len(POS)):
        got = cb[b][iy, ix].astype(np.float64)
        want = ref["samples"][b]
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        tot = abs(float(cb[b].sum(dtype=np.float64)) - float(ref["total"][b])) / float(ref["total"][b])
        print(f"position {POS[b]}: sampled CBED rel-L2 {err:.2e}, total intensity rel. diff {tot:.1e}", flush=True)
        assert err < 1e-4
        assert tot < 1e-4
