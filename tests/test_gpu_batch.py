"""Uneven multi-material batch replayed as a captured hipGraph (MI355X).

Every replay must equal running each material by itself bit for bit (same
kernels, same inputs, same order per material); replays are interleaved with
eager work and synchronisations, and one material is checked against the
FP64 oracle at the 1e-4 CBED bar."""
import sys

import numpy as np
import pytest

from namsa_amd import structures
from namsa_amd.batch import MaterialBatch
from oracle import msa_oracle as O
from tests.parity_util import oracle_pm, oracle_probe, oracle_slices, rel_l2

pytestmark = pytest.mark.gpu


def _specs():
    return structures.multi_material_specs(seed=31, count=6, grids=(128, 192, 256, 384),
                                           thickness=(6.0, 20.0), probes=(1, 6), pix=(0.1, 0.14))


def test_graph_replay_matches_per_material_runs(gpu):
    import torch

    specs = _specs()
    batch = MaterialBatch(specs, streams=3)
    Ns = {m["N"] for m in batch.mats}
    Ss = {m["S"] for m in batch.mats}
    print(f"batch: N={sorted(Ns)} S={sorted(Ss)} B={[m['B'] for m in batch.mats]}", flush=True)
    assert len(Ns) >= 3 and len(Ss) >= 3
    batch.run_sequential()
    torch.cuda.synchronize()
    ref = [m["cbed"].clone() for m in batch.mats]
    batch.capture()
    for rep in range(3):
        for m in batch.mats:
            m["cbed"].fill_(-1.0)
        batch.replay()
        torch.cuda.synchronize()
        for i, m in enumerate(batch.mats):
            assert torch.equal(m["cbed"], ref[i]), (rep, i)
        # eager work between replays: a sequential run and a torch reduction
        batch.run_sequential()
        _ = sum(float(m["cbed"].sum()) for m in batch.mats)
        torch.cuda.synchronize()
    print("replays identical to per-material runs", flush=True)
    sys.stdout.flush()


def test_batch_material_matches_oracle(gpu):
    import torch

    specs = _specs()[:2]
    batch = MaterialBatch(specs, streams=2)
    batch.replay()
    torch.cuda.synchronize()
    for i, m in enumerate(batch.mats):
        sim = m["sim"]
        if m["N"] > 256:
            continue
        t_ref, _ = oracle_slices(sim, specs[i]["dz"])
        psi_k, _ = oracle_probe(sim)
        pm = oracle_pm(sim, specs[i]["dz"], 1 / 3)
        cb_ref = O.cbed(psi_k, sim.probe_positions.astype(np.float32), np.float32(sim.kmax), t_ref, pm)
        err = max(rel_l2(m["cbed"][b].cpu().numpy(), cb_ref[b]) for b in range(m["B"]))
        print(f"material {i} N={m['N']} S={m['S']} B={m['B']} CBED relL2 {err:.2e}", flush=True)
        assert err < 1e-4


def test_batch_large_grids_match_oracle(gpu):
    """512^2, 768^2 and 1024^2 materials inside one captured batch (with a
    256^2 one), replayed; the first probe of each large material against the
    FP64 oracle at the 1e-4 CBED bar."""
    import torch

    specs = []
    for k, N in enumerate((256, 512, 768, 1024)):
        specs += structures.multi_material_specs(seed=70 + k, count=1, grids=(N,), thickness=(6.0, 12.0),
                                                 probes=(2, 3), pix=(0.1, 0.12))
    batch = MaterialBatch(specs, streams=2)
    assert sorted(m["N"] for m in batch.mats) == [256, 512, 768, 1024]
    batch.capture()
    batch.replay()
    batch.replay()
    torch.cuda.synchronize()
    for i, m in enumerate(batch.mats):
        if m["N"] < 512:
            continue
        sim = m["sim"]
        t_ref, _ = oracle_slices(sim, specs[i]["dz"])
        psi_k, _ = oracle_probe(sim)
        pm = oracle_pm(sim, specs[i]["dz"], 1 / 3)
        cb_ref = O.cbed(psi_k, sim.probe_positions[:1].astype(np.float32), np.float32(sim.kmax), t_ref, pm)
        err = rel_l2(m["cbed"][0].cpu().numpy(), cb_ref[0])
        print(f"material {i} N={m['N']} S={m['S']} B={m['B']} CBED relL2 {err:.2e}", flush=True)
        assert err < 1e-4
