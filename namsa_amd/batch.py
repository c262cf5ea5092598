"""Uneven multi-material batches: many materials, one captured HIP graph.

The reference scales over materials by giving every MPI rank its own list of
CIFs and running MSAGPU on each in turn (scripts/summit_scripts/sim_prod.py:157,
:204; scripts/sim_batch.py:106).  On one MI355X the same job is one batch:
every material (its own grid N, slice count S, probe count B, energy and
probe) keeps its slices, probe, propagator and outputs resident, and its
chain of launches -- probe former, S fused slice steps, CBED readout, exactly
the launches namsa_multislice issues for it alone -- is enqueued on one of a
few streams, longest chains first.  The whole batch is captured once into a
hipGraph and replayed; results are bit-identical to running each material by
itself, because every material sees the same kernels on the same inputs in the
same order.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import call, ptr
from .msa import DEFAULT_PROBE, MSAGPU, cache_group


class MaterialBatch:
    """A resident set of materials whose multislice runs as one graph replay.

    specs: dicts with keys records, N, dz, kv, semi (rad), grid_steps, and
    optionally bandwidth (default 1/3) and probe (probe_dict).
    """

    def __init__(self, specs, device=0, streams=4):
        import torch

        self.specs = list(specs)
        self.device = torch.device("cuda", device)
        self.mats = []
        for sp in self.specs:
            N = int(sp["N"])
            sim = MSAGPU(sp["kv"], sp["semi"], sp["records"], sampling=np.array([N, N]),
                         debye_waller=False)
            sim.setup_device(device)
            sim.calc_atomic_potentials()
            sim.build_potential_slices(sp["dz"])
            sim.build_probe(dict(sp.get("probe", DEFAULT_PROBE)))
            sim.generate_probe_positions(grid_steps=np.asarray(sp["grid_steps"]))
            sim.build_propagator(sp.get("bandwidth", 1 / 3))
            B = int(sim.num_probes)
            group = min(B, cache_group(N, N))
            m = {
                "sim": sim, "N": N, "B": B, "S": int(sim.num_slices), "group": group,
                "pos": torch.from_numpy(np.ascontiguousarray(sim.probe_positions, dtype=np.float32)).to(self.device),
                "ws": torch.empty((group, N, N), dtype=torch.complex64, device=self.device),
                "cbed": torch.empty((B, N, N), dtype=torch.float32, device=self.device),
            }
            self.mats.append(m)
        # longest chains first, dealt round-robin over the streams
        self.order = sorted(range(len(self.mats)), key=lambda i: -self.cost(i))
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, streams))]
        self.graph = None

    def cost(self, i):
        m = self.mats[i]
        return m["B"] * m["S"] * m["N"] * m["N"] * np.log2(m["N"])

    @property
    def probe_slices(self):
        return sum(m["B"] * m["S"] for m in self.mats)

    def _launch(self, i, stream):
        m = self.mats[i]
        sim = m["sim"]
        N = m["N"]
        call("namsa_multislice", ptr(sim.psi_k_device), ptr(m["pos"]), m["B"], N, N, np.float32(sim.kmax),
             ptr(sim.potential_slices_device), m["S"], ptr(sim.pm_device), sim.band[0], sim.band[1],
             m["group"], _lib.OUT_CBED, np.float32(1.0 / sim.normalization), ptr(m["ws"]), ptr(m["cbed"]),
             None, 0, None, None, _lib.stream_handle(stream))

    def enqueue(self):
        """Issue every material's chain on the batch streams (fork from / join to the current stream)."""
        import torch

        main = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(main)
        for k, i in enumerate(self.order):
            self._launch(i, self.streams[k % len(self.streams)])
        for s in self.streams:
            main.wait_stream(s)

    def run_sequential(self):
        """Every material in turn on the current stream (the reference's per-material loop)."""
        import torch

        st = torch.cuda.current_stream(self.device)
        for i in range(len(self.mats)):
            self._launch(i, st)

    def capture(self):
        """Record the whole batch once into a hipGraph (after one eager warm-up run)."""
        import torch

        self.enqueue()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.enqueue()
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def cbed(self, i):
        return self.mats[i]["cbed"]
