#!/usr/bin/env python3
"""Benchmark: batched 4D-STEM plain multislice on one MI355X.

Workload (the headline configuration): seeded amorphous-plus-crystal block
64 x 64 x 300 Angstrom (73,728 atoms, Z in {6,14,29,79}), 512 x 512 complex64
grid, dz = 2 Angstrom (150 slices), 200 kV, 25 mrad probe, 64 x 64 raster
(4096 probe positions), float32 CBED per position plus BF (0-25 mrad) and ADF
(40-200 mrad) detector integrals fused into the readout.

A step is one full pass of the hot path over the batch: probe formation,
150 fused slice steps and the CBED/detector readout for all 4096 positions,
with slices, probe and propagator already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line: probe-positions*slices/s, the roofline of the dominant
kernel timed live with HIP events, and a CPU baseline (the op-for-op port of
the reference CPU path) timed on this host's cores for a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "probe-positions·slices/sec at 512² FP32; learned vs plain at ≤1e-3 CBED rel-L2"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--probes", type=int, default=64, help="raster edge (probes = edge^2)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-probes", type=int, default=0, help="CPU sample size (0: 5 per core)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run a 1 GiB device copy first (HBM counter calibration)")
    p.add_argument("--kernel-reps", type=int, default=40)
    p.add_argument("--workload", choices=["config1", "config2", "config3", "config4", "config5", "config5-learned",
                                          "potential3d", "records", "stream"],
                   default="config2",
                   help="config2: 4D-STEM at 512^2 (headline); config3: 48-material uneven batch; "
                        "config4: learned coarse-slice corrector vs plain at 256^2; config5: 2048^2 "
                        "polycrystal, 1000 slices, positions streamed in chunks; potential3d: slices "
                        "integrated from a sampled 3D potential at 512^2")
    p.add_argument("--c5-probes", type=int, default=256,
                   help="config5: raster points timed (evenly spaced over the 64x64 raster; 4096 = all)")
    p.add_argument("--c5-chunk", type=int, default=128, help="config5: positions per streamed chunk")
    p.add_argument("--c5-learned-chunk", type=int, default=6,
                   help="config5-learned: probes per coarse run and corrector call (6: the 2048^2 probe group)")
    p.add_argument("--c4-structures", type=int, default=32, help="config4: held-out structures scored (4 probes each)")
    p.add_argument("--streams", type=int, default=4, help="config3: streams the batch is dealt over")
    p.add_argument("--tune", action="store_true",
                   help="autotune schedules and the probe group first (default: compiled defaults)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(edge):
    import torch

    from namsa_amd import DEFAULT_PROBE, MSAGPU, structures

    rec = structures.random_amorphous_crystal(seed=2)
    sim = MSAGPU(200.0, 0.025, rec, sampling=np.array([512, 512]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(2.0)
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.generate_probe_positions(grid_steps=np.array([edge, edge]))
    sim.build_propagator(1 / 3)
    dets = np.stack([sim.detector_mask(0.0, 25e-3), sim.detector_mask(40e-3, 200e-3)])
    torch.cuda.synchronize()
    return sim, dets


def copy_calibration():
    """Two 1 GiB device-to-device copies (the second timed): the reference
    dispatches that scale the FETCH_SIZE / WRITE_SIZE counters to bytes.
    Returns the copy's read + write rate in GB/s."""
    import torch

    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    b.copy_(a)
    c1.record()
    torch.cuda.synchronize()
    return 2 * a.numel() * 4 / (c0.elapsed_time(c1) * 1e-3) / 1e9


def pass_roofline(sim, pos, ws, G, lo, hi, col_sched, reps):
    """Live per-kernel timing of the two slice-step passes on one group of G
    probes, on the stream the kernels run on: R back-to-back launches of each
    pass between two HIP events (the same launches namsa_multislice issues; a
    pass applied twice in a row costs the same as in alternation).  Returns
    the roofline of the slower (dominant) pass."""
    import torch

    import namsa_amd._lib as L

    Ny, Nx = sim._grid
    plane = Ny * Nx
    S = int(sim.num_slices)
    st = L.stream_handle()
    kmax = np.float32(sim.kmax)
    # the wave layout namsa_multislice's CBED slice loop uses at this size
    # (the probe former's row-major state only seeds finite values)
    sfx = "_tiled" if L.load().namsa_wave_tiled(Ny, Nx) else ""
    pm = sim.cbed_propagator()[0] if sfx else L.ptr(sim.pm_device)
    L.call("namsa_probe_init", L.ptr(sim.psi_k_device), L.ptr(pos), G, Ny, Nx, kmax, L.ptr(ws), st)
    L.call("namsa_slice_rows" + sfx, L.ptr(ws), G, Ny, Nx, L.ptr(sim._t[0]), 0, Nx, st)
    L.call("namsa_slice_cols" + sfx, L.ptr(ws), G, Ny, Nx, pm, lo, hi, st)
    reps = max(2, reps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for r in range(reps):
        L.call("namsa_slice_rows" + sfx, L.ptr(ws), G, Ny, Nx, L.ptr(sim._t[1 + r % max(1, S - 1)]), lo, hi, st)
    ev[1].record()
    ev[2].record()
    for r in range(reps):
        L.call("namsa_slice_cols" + sfx, L.ptr(ws), G, Ny, Nx, pm, lo, hi, st)
    ev[3].record()
    torch.cuda.synchronize()
    row_ms = ev[0].elapsed_time(ev[1]) / reps
    col_ms = ev[2].elapsed_time(ev[3]) / reps
    frac = (hi - lo) / Nx
    strip = 16 if "c16" in col_sched else (8 if "c8" in col_sched else 4)
    strips = (-(-hi // strip) - lo // strip) * strip / Nx
    # Algorithmic bytes per probe-slice: 32 N^2 + 8 N^2 / B (two read+write
    # round trips of the complex64 wave, the slice read once per launch):
    # 16 N^2 + 8 N^2 / B for the row pass, 16 N^2 for the column pass.
    row_alg = 16 * plane * G + 8 * plane
    col_alg = 16 * plane * G
    # Bytes the band-limited kernels actually move (loads of the open band,
    # stores of the 16-aligned band / band strips).
    sfrac = (min(Nx, -(-hi // 16) * 16) - lo // 16 * 16) / Nx
    row_band = int((8 * frac + 8 * sfrac) * plane * G + 8 * plane)
    col_band = int(16 * strips * plane * G)
    kernels = {"row_pass": (row_ms, row_alg, row_band), "col_pass": (col_ms, col_alg, col_band)}
    dom = max(kernels, key=lambda k: kernels[k][0])
    d_ms, d_bytes, d_band = kernels[dom]
    achieved = d_bytes / (d_ms * 1e-3) / 1e9
    step_gbs = (row_alg + col_alg) / ((row_ms + col_ms) * 1e-3) / 1e9
    step_band = (row_band + col_band) / ((row_ms + col_ms) * 1e-3) / 1e9
    log(f"group {G}: row pass {row_ms * 1e3:.1f} us ({row_alg / row_ms / 1e6:.0f} GB/s algorithmic, "
        f"{row_band / row_ms / 1e6:.0f} GB/s moved), col pass {col_ms * 1e3:.1f} us ({col_alg / col_ms / 1e6:.0f} "
        f"GB/s algorithmic, {col_band / col_ms / 1e6:.0f} GB/s moved), slice step {step_gbs:.0f} / {step_band:.0f} GB/s")
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": d_bytes, "avg_launch_ms": d_ms,
            "launch_probes": G, "band_bytes_per_launch": d_band, "band_GBs": d_band / (d_ms * 1e-3) / 1e9,
            "band_frac": d_band / (d_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "slice_step_GBs": step_gbs,
            "slice_step_band_GBs": step_band, "row_ms": row_ms, "col_ms": col_ms,
            "wave_layout": "2x4 tiles" if sfx else "row-major"}


def host_cpu_info():
    """The host the CPU baseline runs on: core counts, affinity, model, numpy."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"os_cpu_count": os.cpu_count(), "affinity": affinity, "model": model, "numpy": np.__version__}


# Worker processes of the CPU baseline: the reference sizes its Pool as
# min(os.cpu_count(), probes) (msa.py:295); one GPU's share of the box's host
# is 16 cores, whatever os.cpu_count() reports there, so the pool is capped at that.
CPU_SHARE = 16


def cpu_config1():
    """Config 1 in full on the CPU path: Si[111], 256^2, dz = 1.92 A, one probe
    at the centre, 100 kV, 20 mrad (oracle.cpu_port: the numpy branch of
    msa.py:280-361, plus the CBED readout of the exit wave), one process."""
    from namsa_amd import MSAGPU, structures
    from oracle import cpu_port
    from oracle import msa_oracle as O

    rec = structures.si111_supercell()
    sim = MSAGPU(100.0, 0.020, rec, sampling=np.array([256, 256]), debye_waller=False)
    kirk = np.load(ROOT / "namsa_amd" / "data" / "kirkland_params.npy")
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, 1.92, sim.pix_size, sim.dims)
    tables = np.stack([O.atomic_table(kirk[z - 1], sim.pix_size) for z in uniq])
    V = O.splat_slices(tables, sites, 256, 256).astype(np.float32)
    t0 = time.perf_counter()
    cb = cpu_port.multislice_cpu([(0.0, 0.0)], V, sim.sigma, sim.kmax, sim.Lambda, 1.92, sim.semi_ang,
                                 bandwidth=1 / 3, probe_kw=dict(apert_smooth=80.0), processes=1, cbed=True)
    dt = time.perf_counter() - t0
    S = V.shape[0]
    return {"value": S / dt, "unit": "probe-positions·slices/s", "wall_s": dt, "cores": 1, "kind": "port",
            "finite": bool(np.isfinite(cb).all()),
            "sample": f"config 1 in full: Si[111] {len(rec)} atoms, 1 probe x {S} slices at 256^2 + CBED readout"}


def cpu_baseline(sim, n_probes):
    """The reference CPU path (oracle.cpu_port, with the CBED readout) on
    n_probes raster points of the 512^2 workload, plus config 1 in full."""
    from oracle import cpu_port
    from oracle import msa_oracle as O

    kirk = np.load(ROOT / "namsa_amd" / "data" / "kirkland_params.npy")
    sites, uniq = O.bin_sites(sim.supercell_xyz, sim.supercell_Z, sim.slice_t, sim.pix_size, sim.dims)
    tables = np.stack([O.atomic_table(kirk[z - 1], sim.pix_size) for z in uniq])
    V = O.splat_slices(tables, sites, 512, 512).astype(np.float32)
    host = host_cpu_info()
    cores = max(1, min(host["affinity"], CPU_SHARE))
    n = n_probes or 5 * cores                  # five points per core: about 10 s of CPU work
    idx = np.linspace(0, len(sim.probe_positions) - 1, n).astype(int)
    # the CPU path takes positions as (x, y) (msa.py:225)
    pos = sim.probe_positions[idx][:, ::-1]
    t0 = time.perf_counter()
    cb = cpu_port.multislice_cpu(pos, V, sim.sigma, sim.kmax, sim.Lambda, sim.slice_t, sim.semi_ang,
                                 bandwidth=1 / 3, probe_kw=dict(apert_smooth=80.0), processes=min(cores, n),
                                 cbed=True)
    dt = time.perf_counter() - t0
    S = V.shape[0]
    return {"value": n * S / dt, "unit": "probe-positions·slices/s", "cores": min(cores, n),
            "kind": "port", "host": host, "finite": bool(np.isfinite(cb).all()),
            "sample": f"{n} raster points x {S} slices at 512^2 with the CBED readout (oracle/cpu_port.py, "
                      f"numpy {np.__version__}, fork Pool of {min(cores, n)}: the reference's "
                      f"min(cpu_count, probes), capped at the {CPU_SHARE}-core share of one GPU), {dt:.1f} s",
            "config1": cpu_config1()}


def bench_config1(args):
    """Config 1: Si[111] 38 x 39 x 200 A (15,360 atoms), 256^2, dz = 1.92 A
    (104 slices), 100 kV, 20 mrad, one probe at the centre -- the reference's
    CPU-path workload -- through the fused GPU driver, beside the CPU path
    timed in full."""
    import torch

    from namsa_amd import DEFAULT_PROBE, MSAGPU, structures

    sim = MSAGPU(100.0, 0.020, structures.si111_supercell(), sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(1.92)
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.probe_positions = np.zeros((1, 2))
    sim.num_probes = np.int32(1)
    sim.grid_steps = np.array([1, 1])
    sim.plan_simulation()
    S = int(sim.num_slices)
    for _ in range(max(1, args.warmup)):
        sim.multislice(bandwidth=1 / 3)
    reps = max(5, args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.multislice(bandwidth=1 / 3)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    total = float(sim.probes.sum())
    cpu = None if args.no_cpu_baseline else cpu_config1()
    out = {
        "metric": "probe-positions·slices/s, config 1 (single probe, CPU-path workload)",
        "value": S / wall, "unit": "probe-positions·slices/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (Si[111] supercell, deterministic)",
        "config": {"workload": "config 1: Si[111] 256^2, 104 slices, 1 probe, CBED (drop-in multislice call, "
                               "host overheads included)", "grid": [256, 256], "slices": S, "probes": 1},
        "cbed_sum": total, "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_potential3d(args):
    """Transmission slices from a sampled 3D potential (integrate_potential_slices):
    a 512^2 plane, 1200 z samples (FP64), 150 slices of 8 samples, timed on the
    kernel's stream; the CPU baseline is the FP64 restatement of the
    reference's numpy path (crop, split, trapz, exp) on a bounded sample."""
    import torch

    import namsa_amd._lib as L

    N, S, d = 512, 150, 8
    Z = S * d
    rng = np.random.default_rng(7)
    dev = torch.device("cuda", 0)
    z = np.arange(Z) * 0.25
    # smooth synthetic field, generated on the device (2.5 GB of FP64 samples)
    zz = torch.from_numpy(z).to(dev)[:, None, None]
    yy = torch.arange(N, device=dev, dtype=torch.float64)[None, :, None] * 0.1
    xx = torch.arange(N, device=dev, dtype=torch.float64)[None, None, :] * 0.1
    kz, ky, kx = rng.uniform(0.5, 2.0, 3)
    pot = -20.0 * (1.5 + torch.cos(kz * zz) * torch.cos(ky * yy) * torch.cos(kx * xx))
    pot = pot.contiguous()
    z_d = torch.from_numpy(z).to(dev)
    t = torch.empty((S, N, N), dtype=torch.complex64, device=dev)
    sigma = 7.2884e-4
    st = L.stream_handle()

    def step():
        L.call("namsa_integrate_slices", L.ptr(pot), L.ptr(z_d), S, d, N, N, sigma, L.ptr(t), None, st)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.steps) * 4
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 8 * Z * N * N + 8 * Z + 8 * S * N * N       # samples + coordinates read, slices written
    gbs = nbytes / (ms * 1e-3) / 1e9
    value = Z * N * N / (ms * 1e-3)
    log(f"potential3d: {Z} x {N}^2 FP64 samples -> {S} slices in {ms:.3f} ms, {gbs:.0f} GB/s")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import msa_oracle as O

        host = pot[: 20 * d + 1].cpu().numpy()
        grid_h = (z[: 20 * d + 1], np.arange(N) * 0.1, np.arange(N) * 0.1)
        passes, t0 = 0, time.perf_counter()
        while True:                      # repeat the 20-slice sample for about 10 s of CPU work
            V, info = O.integrate_slices(host, grid_h, 0.25 * d)
            np.exp(1j * sigma * V).astype(np.complex64)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= 10.0:
                break
        cpu = {"value": passes * info["z_used"] * N * N / dt, "unit": "samples/s", "cores": 1, "kind": "port",
               "sample": f"{passes} passes over {info['num_slices']} slices of {info['d']} samples at {N}^2 "
                         f"(oracle/msa_oracle.py integrate_slices + exp, numpy {np.__version__}, one process), "
                         f"{dt:.1f} s"}
        log(f"cpu baseline {cpu['value']:.4g} samples/s")
    out = {
        "metric": "potential samples integrated into transmission slices per second", "value": value,
        "unit": "samples/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (smooth periodic field)",
        "config": {"workload": f"integrate_potential_slices: {N}x{N} plane, {Z} z samples (FP64), {S} slices "
                               f"of {d} samples, exp(i sigma V) to complex64", "grid": [N, N], "samples": Z,
                   "slices": S},
        "roofline": {"bound": "hbm", "kernel": "integrate_slices_kernel", "achieved": gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": _aux_traffic("potential3d"),
                     "bytes_per_launch": nbytes,
                     "avg_launch_ms": ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_stream(args):
    """The 512^2 workload (4096 positions, 150 slices) with its CBEDs delivered
    to host memory: compute only; compute then a blocking copy to pinned
    memory; and multislice(sink=...), which copies each finished chunk on a
    second stream while the next chunk computes. The same for the float16
    records (readout moments, one-pass encoder)."""
    import torch

    from namsa_amd import records as REC

    sim, _ = setup(64)
    sim.plan_simulation()
    P, S = int(sim.num_probes), int(sim.num_slices)
    delivered = [0]

    def sink(start, arr):
        delivered[0] += len(arr)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    pinned = torch.empty((P, 512, 512), dtype=torch.float32, pin_memory=True)
    pinned16 = torch.empty((P, 512, 512), dtype=torch.float16, pin_memory=True)

    def blocking():
        sim.multislice(bandwidth=1 / 3)
        pinned.copy_(sim.probes_device)

    def blocking_records():
        sim.multislice(bandwidth=1 / 3, moments=True)
        pinned16.copy_(REC.process_cbed(sim.probes_device, moments=sim.cbed_moments_device, as_tensor=True))

    for _ in range(max(1, args.warmup)):
        sim.multislice(bandwidth=1 / 3, sink=sink)
    reps = max(1, min(args.steps, 3))
    t = {"compute_only": [], "blocking_copy": [], "streamed": [], "records_blocking": [], "records_streamed": []}
    for _ in range(reps):
        t["compute_only"].append(timed(lambda: sim.multislice(bandwidth=1 / 3)))
        t["blocking_copy"].append(timed(blocking))
        delivered[0] = 0
        t["streamed"].append(timed(lambda: sim.multislice(bandwidth=1 / 3, sink=sink)))
        assert delivered[0] == P
        t["records_blocking"].append(timed(blocking_records))
        delivered[0] = 0
        t["records_streamed"].append(timed(lambda: sim.multislice(bandwidth=1 / 3, sink=sink, records=True)))
        assert delivered[0] == P
    ms = {k: float(np.median(v)) for k, v in t.items()}
    log("stream: " + ", ".join(f"{k} {v:.1f} ms" for k, v in ms.items()))
    out = {
        "metric": "probe-positions·slices/s at 512² with every CBED delivered to host memory",
        "value": P * S / (ms["streamed"] * 1e-3), "unit": "probe-positions·slices/s", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "ms_per_step": ms["streamed"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded amorphous block, 73,728 atoms)",
        "config": {"workload": "config-2 multislice (4096 positions, 150 slices, 512^2) with CBEDs streamed to "
                               "pinned host memory in chunks", "grid": [512, 512], "probes": P, "slices": S},
        "timings_ms": ms,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def _aux_traffic(workload):
    """HBM bytes per call measured with the FETCH_SIZE / WRITE_SIZE counters
    (profiles/aux_workloads_hbm_traffic.txt), or None."""
    tf = ROOT / "profiles" / "hbm_traffic_aux.json"
    if not tf.exists():
        return None
    return json.loads(tf.read_text()).get(workload, {}).get("traffic")


def bench_records(args):
    """Training records of a 4D-STEM dataset (process_cbed): 4096 CBEDs of
    512^2 float32 already in HBM, standardised per pattern to half. One call
    is one step; the CPU baseline is the oracle's numpy float32 restatement of
    the reference's process_cbed on a bounded sample."""
    import torch

    import namsa_amd._lib as L

    P, N = 4096, 512
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    cb = (torch.rand((P, N, N), generator=g, device=dev) ** 6).contiguous()     # heavy-tailed, nonnegative
    out = torch.empty((P, N, N), dtype=torch.float16, device=dev)
    lib = L.load()
    ws = torch.empty(int(lib.namsa_records_workspace_doubles(P, N, N)), dtype=torch.float64, device=dev)
    st = L.stream_handle()

    def step():
        L.call("namsa_cbed_records", L.ptr(cb), P, N, N, L.ptr(out), L.ptr(ws), st)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.steps)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 6 * P * N * N                      # one float32 read and one half write per pixel
    gbs = nbytes / (ms * 1e-3) / 1e9
    value = P / (ms * 1e-3)
    log(f"records: {P} x {N}^2 CBEDs -> half in {ms:.3f} ms, {gbs:.0f} GB/s")
    # The same records when the CBED readout kept each pattern's sums (multislice(moments=True)):
    # one pass. The sums are formed here, before the timed region, in 256-pattern slabs.
    mom = torch.empty((P, 2), dtype=torch.float64, device=dev)
    for p0 in range(0, P, 256):
        slab = cb[p0:p0 + 256].double()
        mom[p0:p0 + 256, 0] = slab.sum(dim=(1, 2))
        mom[p0:p0 + 256, 1] = (slab * slab).sum(dim=(1, 2))
        del slab
    out1 = torch.empty_like(out)
    L.call("namsa_cbed_records_from_moments", L.ptr(cb), L.ptr(mom), P, N, N, L.ptr(out1), st)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        L.call("namsa_cbed_records_from_moments", L.ptr(cb), L.ptr(mom), P, N, N, L.ptr(out1), st)
    e1.record()
    torch.cuda.synchronize()
    ms1 = e0.elapsed_time(e1) / reps
    same = float((out1.view(torch.int16) == out.view(torch.int16)).double().mean())
    log(f"records from readout moments: {ms1:.3f} ms, {nbytes / (ms1 * 1e-3) / 1e9:.0f} GB/s, "
        f"{100 * same:.3f}% identical to the two-pass records")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import records_oracle as R

        host = cb[:64].cpu().numpy()
        passes, t0 = 0, time.perf_counter()
        while True:                              # about 10 s of CPU work
            R.process_cbed(host)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= 10.0:
                break
        cpu = {"value": passes * 64 / dt, "unit": "CBEDs/s", "cores": 1, "kind": "port",
               "sample": f"{passes} passes over 64 CBEDs of {N}^2 (oracle/records_oracle.py process_cbed, numpy "
                         f"{np.__version__}, one process), {dt:.1f} s"}
        log(f"cpu baseline {cpu['value']:.4g} CBEDs/s")
    out_line = {
        "metric": "CBEDs encoded into standardised half records per second", "value": value,
        "unit": "CBEDs/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (heavy-tailed uniform^6 patterns)",
        "config": {"workload": f"process_cbed: {P} CBEDs of {N}x{N} float32 -> float16", "probes": P,
                   "grid": [N, N]},
        "roofline": {"bound": "hbm", "kernel": "rec_stats_kernel + rec_standardize_kernel", "achieved": gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": _aux_traffic("records"), "bytes_per_launch": nbytes, "avg_launch_ms": ms,
                     "from_readout_moments": {"kernel": "rec_standardize_kernel", "avg_launch_ms": ms1,
                                              "achieved": nbytes / (ms1 * 1e-3) / 1e9,
                                              "frac": nbytes / (ms1 * 1e-3) / 1e9 / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out_line), flush=True)


def bench_config3(args):
    """48 seeded random crystals, uneven grids / slices / probes, one graph replay."""
    import torch

    from namsa_amd import structures
    from namsa_amd.batch import MaterialBatch

    log("config3: generating 48 crystals and building their slices, probes, propagators")
    specs = structures.multi_material_specs(seed=3)
    batch = MaterialBatch(specs, streams=args.streams)
    ps = batch.probe_slices
    torch.cuda.synchronize()
    times = {}
    for mode in ("sequential", "streams", "graph"):
        fn = {"sequential": batch.run_sequential, "streams": batch.enqueue, "graph": batch.replay}[mode]
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        times[mode] = (time.perf_counter() - t0) / args.steps
        log(f"{mode}: {times[mode] * 1e3:.1f} ms per batch, {ps / times[mode]:.4g} probe-slices/s")
    grids = sorted({m["N"] for m in batch.mats})
    # algorithmic HBM bytes of the batch: per material S (32 N^2 B + 8 N^2 ceil(B / group)) -- the
    # complex64 wave read and written twice per probe-slice, each slice read once per probe group
    alg = sum(m["S"] * (32 * m["N"] ** 2 * m["B"] + 8 * m["N"] ** 2 * -(-m["B"] // m["group"])) for m in batch.mats)
    achieved = alg / times["graph"] / 1e9
    per_size = {}
    for m in batch.mats:
        e = per_size.setdefault(m["N"], [0, 0])
        e[0] += 1
        e[1] += m["B"] * m["S"]
    log(f"config3 roofline: {alg / 1e9:.1f} GB algorithmic per replay, {achieved:.0f} GB/s "
        f"({achieved / HBM_PEAK_GBS:.2f} of {HBM_PEAK_GBS:.0f})")
    # A/B against launching each material alone: per grid, the time its materials take one at a
    # time (HIP events) and the rate on their algorithmic bytes
    st = torch.cuda.current_stream()
    alone = {}
    for i, m in enumerate(batch.mats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        batch._launch(i, st)
        e0.record(st)
        batch._launch(i, st)
        e1.record(st)
        torch.cuda.synchronize()
        b = m["S"] * (32 * m["N"] ** 2 * m["B"] + 8 * m["N"] ** 2 * -(-m["B"] // m["group"]))
        a = alone.setdefault(m["N"], [0.0, 0.0, 0])
        a[0] += e0.elapsed_time(e1) * 1e-3
        a[1] += b
        a[2] += m["B"] * m["S"]
    per_grid = {str(N): {"seconds_alone": v[0], "GBs_alone": v[1] / v[0] / 1e9, "probe_slices": v[2]}
                for N, v in sorted(alone.items())}
    t_alone = sum(v[0] for v in alone.values())
    log("config3 per grid, each material alone: " + ", ".join(
        f"{N}: {v['seconds_alone'] * 1e3:.0f} ms at {v['GBs_alone']:.0f} GB/s" for N, v in per_grid.items()))
    log(f"config3: materials alone sum to {t_alone * 1e3:.0f} ms; one graph replay {times['graph'] * 1e3:.0f} ms")
    out = {
        "metric": METRIC, "value": ps / times["graph"], "unit": "probe-positions·slices/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": times["graph"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (48 seeded random crystals)",
        "config": {"workload": "uneven multi-material batch: 48 seeded cubic/tetragonal/hexagonal crystals, "
                               "a in [3,7] A, 1-8 basis atoms, Z in [3,83], 50-500 A thick, grids "
                               f"{grids}, dz 1-3 A, 60-200 kV, 10-30 mrad, 1-256 probes each, "
                               "one captured hipGraph",
                   "materials": len(batch.mats), "probe_slices": ps, "streams": args.streams,
                   "sequential_probe_slices_per_s": ps / times["sequential"],
                   "multistream_probe_slices_per_s": ps / times["streams"],
                   "materials_and_probe_slices_per_grid": {str(k): v for k, v in sorted(per_size.items())},
                   "each_material_alone_per_grid": per_grid, "each_material_alone_total_s": t_alone},
        "roofline": {"bound": "hbm", "kernel": "whole batch (row and column passes of every material)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_replay": alg,
                     "accounting": "sum over materials of S (32 N^2 B + 8 N^2 ceil(B / group))"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_config4(args):
    """Learned coarse-slice corrector vs plain fine multislice at 256^2 (the
    corrector's training configuration): seeded random crystals 32 A wide,
    200 kV, 25 mrad, coarse dz = f A plus one spectral correction per probe
    (FNO, 4 layers, 32 channels, 16x16 modes; weights trained by
    namsa_amd.train_corrector on 18k pairs, committed under namsa_amd/data)
    against fine dz = 1 A.  Accuracy: CBED rel-L2 against the fine path on
    held-out structures from a third seed range.  Throughput: 64 probes at
    200 and 400 A.  Every factor whose weights exist forms the error-vs-cost
    curve."""
    import torch

    from namsa_amd import corrector as K

    data = ROOT / "namsa_amd" / "data"
    nets = {}
    for f in (2, 4, 8):
        path = data / f"corrector_c4_f{f}.npz"
        if path.exists():
            nets[f] = K.SpectralCorrector.load(str(path))
    if not nets:
        raise SystemExit("no corrector weights under namsa_amd/data; train them with python -m namsa_amd.train_corrector")
    factors = tuple(sorted(nets))
    errs = {f: {"coarse": [], "learned": []} for f in factors}
    above = {f: 0 for f in factors}          # probes the default fallback threshold would recompute
    t0 = time.perf_counter()
    for seed in range(950_000, 950_000 + args.c4_structures):
        sim, rng, _ = K.random_crystal_sim(seed)
        mat = K.SpectralMaterial(sim, factors=factors)
        S = mat.S_f // 8 * 8
        L = float(min(sim.dims[:2]))
        pos = torch.from_numpy(rng.uniform(-L / 4, L / 4, (4, 2)).astype(np.float32)).to(sim.device)
        fine = mat.cbed(pos, S, None, torch.empty((4, 256, 256), device=sim.device)).cpu().numpy()
        for f in factors:
            coarse = mat.cbed(pos, S // f, f, torch.empty((4, 256, 256), device=sim.device)).cpu().numpy()
            st = {}
            learned = K.run_spectral(nets[f], mat, pos, float(S), tau=1e30, stats=st).cpu().numpy()
            above[f] += int((st["estimate"] > K.DEFAULT_TAU).sum())
            for b in range(4):
                n = np.linalg.norm(fine[b])
                errs[f]["coarse"].append(float(np.linalg.norm(coarse[b] - fine[b]) / n))
                errs[f]["learned"].append(float(np.linalg.norm(learned[b] - fine[b]) / n))
    log(f"config4: {args.c4_structures} held-out structures scored in {time.perf_counter() - t0:.1f} s")

    def rate(fn, B):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return B * args.steps / (time.perf_counter() - t1)

    # throughput on one held-out crystal 410 A thick (at least 400 fine slices whatever its lattice)
    sim, rng, _ = K.random_crystal_sim(960_000, thickness=(410.0, 410.0))
    mat = K.SpectralMaterial(sim, factors=factors)
    if mat.S_f < 400:
        raise SystemExit(f"throughput structure has {mat.S_f} fine slices, fewer than 400")
    B = 64
    L = float(min(sim.dims[:2]))
    pos = torch.from_numpy(rng.uniform(-L / 4, L / 4, (B, 2)).astype(np.float32)).to(sim.device)
    out = torch.empty((B, 256, 256), device=sim.device)
    curve = {}
    for T in (200, 400):
        row = {"fine_probe_positions_per_s": rate(lambda: mat.cbed(pos, T, None, out), B)}
        for f in factors:
            vh = mat.potential_spectrum(T, torch.empty((256, 256), dtype=torch.complex64, device=sim.device))
            row[f"coarse_f{f}"] = rate(lambda: mat.cbed(pos, T // f, f, out), B)
            row[f"learned_f{f}"] = rate(lambda: K.run_spectral(nets[f], mat, pos, float(T), vhat=vh, tau=None), B)
        curve[f"{T}A"] = row
        log(f"config4 throughput at {T} A (probe-positions/s): " + ", ".join(f"{k} {v:.0f}" for k, v in row.items()))
    # the corrector's forward pass (channel-last BF16 kernels) on the same batch, timed with events on its stream
    f = max(factors)
    net = nets[f].net
    t_fwd, tflops, gbs = corrector_forward_rate(net, B)
    summary = {f: {"coarse_mean": float(np.mean(errs[f]["coarse"])), "learned_mean": float(np.mean(errs[f]["learned"])),
                   "coarse_median": float(np.median(errs[f]["coarse"])),
                   "learned_median": float(np.median(errs[f]["learned"])),
                   "fraction_improved": float(np.mean(np.array(errs[f]["learned"]) < np.array(errs[f]["coarse"]))),
                   "probes_above_default_fallback_tau": above[f]}
               for f in factors}
    for f in factors:
        log(f"config4 factor {f}: CBED rel-L2 vs fine, coarse mean {summary[f]['coarse_mean']:.4f} -> learned "
            f"{summary[f]['learned_mean']:.4f} (median {summary[f]['coarse_median']:.4f} -> "
            f"{summary[f]['learned_median']:.4f}), improved in {summary[f]['fraction_improved']:.0%}")
    r400 = curve["400A"]
    ratio = r400[f"learned_f{f}"] / r400["fine_probe_positions_per_s"]
    out = {
        "metric": METRIC, "value": r400[f"learned_f{f}"], "unit": "probe-positions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": B / r400[f"learned_f{f}"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 MFMA (corrector), fp32",
        "data": "synthetic (seeded random crystals; corrector weights trained on 18k seeded pairs)",
        "config": {"workload": f"learned coarse multislice (dz = {f} A + one spectral correction per probe) vs plain "
                               "fine (dz = 1 A), 256^2, 200 kV, 25 mrad, 64 probes, 400 A",
                   "learned_over_plain_fine_400A": ratio,
                   "learned_over_plain_fine_200A": curve["200A"][f"learned_f{f}"] / curve["200A"][
                       "fine_probe_positions_per_s"],
                   "cbed_relL2_vs_fine": {str(k): v for k, v in summary.items()},
                   "held_out_structures": args.c4_structures, "probes_per_structure": 4,
                   "throughput_probe_positions_per_s": curve},
        "roofline": corrector_roofline(net, B, t_fwd, tflops, gbs),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def corrector_forward_rate(net, B, reps=5):
    """Seconds per forward_fused pass on B samples (HIP events on the corrector's
    stream), its TFLOP/s and its algorithmic GB/s."""
    import torch

    net.ensure_fused(B)
    net.forward_fused(B)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        net.forward_fused(B)
    ev1.record(stream)
    torch.cuda.synchronize()
    t = ev0.elapsed_time(ev1) / reps * 1e-3
    return t, net.flops(B) / t / 1e12, net.fused_bytes(B) / t / 1e9


def corrector_roofline(net, B, t, tflops, gbs):
    """The corrector forward is HBM-bound (about 30 kflop per pixel against
    about 0.6 KB of bf16 activation traffic, below the 2.5 PFLOP/s / 8 TB/s
    ridge): the HBM rate on its algorithmic bytes, with the MFMA rate beside it."""
    return {"bound": "hbm", "kernel": "corrector forward (channel-last BF16 MFMA kernels + mode-space GEMMs)",
            "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_pass": net.fused_bytes(B), "forward_ms": t * 1e3, "batch": B,
            "mfma_TFLOPs": tflops, "mfma_frac": tflops / 2500.0, "flops_per_pass": net.flops(B)}


def bench_config5(args):
    """Seeded 4-grain polycrystal 128 x 128 x 1000 A at 2048^2, dz = 1 A (1000
    slices), 300 kV, 30 mrad; raster positions streamed through the GPU in
    chunks with one reused CBED buffer, BF / ADF / total-intensity signals
    kept for every position."""
    import torch

    import namsa_amd._lib as L
    from namsa_amd import DEFAULT_PROBE, MSAGPU, structures
    from namsa_amd import autotune
    from namsa_amd.msa import cache_group

    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    rec, info = structures.polycrystal(seed=5)
    log(f"config5: polycrystal {info['atoms']} atoms, grains {info['grains']} ({time.perf_counter() - t0:.1f} s)")
    sim = MSAGPU(300.0, 0.030, rec, sampling=np.array([2048, 2048]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_potential_slices(1.0)
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.generate_probe_positions(grid_steps=np.array([64, 64]))
    sim.build_propagator(1 / 3)
    torch.cuda.synchronize()
    log(f"config5: {int(sim.num_slices)} slices at 2048^2 built ({time.perf_counter() - t0:.1f} s)")
    Ny, Nx = sim._grid
    S = int(sim.num_slices)
    dev = sim.device
    P = max(1, min(args.c5_probes, len(sim.probe_positions)))
    idx = np.linspace(0, len(sim.probe_positions) - 1, P).astype(int)
    pos = torch.from_numpy(np.ascontiguousarray(sim.probe_positions[idx], dtype=np.float32)).to(dev)
    chunk = max(1, min(args.c5_chunk, P))
    lo, hi = sim.band
    group = min(cache_group(Ny, Nx), chunk)
    dets = np.stack([sim.detector_mask(0.0, 30e-3), sim.detector_mask(40e-3, 200e-3),
                     np.ones((Ny, Nx), np.float32)])
    det = torch.from_numpy(dets).to(dev)
    lib = L.load()
    det_ws = torch.empty(int(lib.namsa_det_workspace_doubles(chunk, Ny, Nx, 3)), dtype=torch.float64, device=dev)
    det_out = torch.empty((P, 3), dtype=torch.float64, device=dev)
    cbed = torch.empty((chunk, Ny, Nx), dtype=torch.float32, device=dev)
    ws = torch.empty((group, Ny, Nx), dtype=torch.complex64, device=dev)
    scale = np.float32(1.0 / sim.normalization)
    kmax = np.float32(sim.kmax)
    st = L.stream_handle()

    pm_arg, pm_flags = sim.cbed_propagator()

    def step():
        for c0 in range(0, P, chunk):
            B = min(chunk, P - c0)
            L.call("namsa_multislice", L.ptr(sim.psi_k_device), L.ptr(pos[c0]), B, Ny, Nx, kmax, L.ptr(sim._t), S,
                   pm_arg, lo, hi, group, L.OUT_CBED | pm_flags, scale, L.ptr(ws), L.ptr(cbed), L.ptr(det),
                   3, L.ptr(det_ws), L.ptr(det_out[c0]), st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    value = P * S * args.steps / wall
    sig = det_out.cpu().numpy()
    log(f"config5: {P} positions x {S} slices in chunks of {chunk}, groups of {group}: {wall / args.steps:.2f} s "
        f"per step, {value:.4g} probe-slices/s; total intensity {sig[:, 2].min():.4f}..{sig[:, 2].max():.4f}")
    # time the schedules the driver ran (thick runs use the full-twiddle forms)
    keep = (autotune.selected(Nx, 0), autotune.selected(Nx, 1))
    run_row, run_col = autotune.driver_schedules(Nx, S)
    autotune.select(Nx, 0, run_row)
    autotune.select(Nx, 1, run_col)
    sched = {"row": autotune.schedules(Nx, 0)[run_row], "col": autotune.schedules(Nx, 1)[run_col],
             "group": group, "band": [lo, hi]}
    rf = pass_roofline(sim, pos, ws, group, lo, hi, sched["col"], args.kernel_reps)
    autotune.select(Nx, 0, keep[0])
    autotune.select(Nx, 1, keep[1])
    tf = ROOT / "profiles" / "hbm_traffic_config5_2048.json"
    if tf.exists():
        rf["traffic"] = json.loads(tf.read_text()).get(rf["kernel"])
    out = {
        "metric": METRIC, "value": value, "unit": "probe-positions·slices/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded polycrystal)",
        "config": {"workload": "2048^2 plain multislice: seeded 4-grain Voronoi polycrystal (Si/GaAs/Au-like "
                               "lattices, random 3D orientations) 128x128x1000 A, dz = 1 A (1000 slices), 300 kV, "
                               f"30 mrad, {P} of the 64x64 raster positions streamed in chunks of {chunk}, "
                               "BF / ADF / total-intensity signals",
                   "atoms": info["atoms"], "grid": [Ny, Nx], "slices": S, "probe_positions": P, "chunk": chunk,
                   "method": "plain", "schedules": sched,
                   "total_intensity_range": [float(sig[:, 2].min()), float(sig[:, 2].max())]},
        "roofline": rf, "cpu_baseline": None, "finite_outputs": bool(np.isfinite(sig).all()),
    }
    print(json.dumps(out), flush=True)


def bench_config5_learned(args):
    """Config 5 in its two forms on the seed-5 polycrystal (128 x 128 x 1000 A,
    2048^2, 300 kV, 30 mrad): plain fine multislice (dz = 1 A, 1000 slices)
    against learned coarse multislice (dz = 8 A, 125 slices, plus one spectral
    correction per probe by the network trained on polycrystals of other
    seeds, namsa_amd/data/corrector_poly_f8.npz).  CBED rel-L2 on 64 positions
    evenly spread over the 64 x 64 raster, against the fine GPU path (itself
    within 3.8e-5 of the FP64 oracle at this size); probe-positions/s of
    both forms and of the plain coarse path."""
    import torch

    from namsa_amd import DEFAULT_PROBE, MSAGPU, structures
    from namsa_amd import corrector as K
    from namsa_amd.msa import cache_group

    path = ROOT / "namsa_amd" / "data" / "corrector_poly_f8.npz"
    if not path.exists():
        raise SystemExit(f"{path} missing; train it with python -m namsa_amd.train_corrector --preset poly")
    corr = K.SpectralCorrector.load(str(path))
    t0 = time.perf_counter()
    rec, info = structures.polycrystal(seed=5)
    sim = MSAGPU(300.0, 0.030, rec, sampling=np.array([2048, 2048]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    sim.build_probe(dict(DEFAULT_PROBE))
    sim.generate_probe_positions(grid_steps=np.array([64, 64]))
    mat = K.SpectralMaterial(sim, dz=1.0, factors=(8,))
    torch.cuda.synchronize()
    log(f"config5 learned: {info['atoms']} atoms, {mat.S_f} fine / {mat.S_c[8]} coarse slices at 2048^2 "
        f"({time.perf_counter() - t0:.1f} s)")
    dev = sim.device
    P = 64
    idx = np.linspace(0, len(sim.probe_positions) - 1, P).astype(int)
    pos = torch.from_numpy(np.ascontiguousarray(sim.probe_positions[idx], dtype=np.float32)).to(dev)
    T = float(mat.S_c[8] * 8)
    S_f = int(T)
    G = cache_group(2048, 2048)
    chunk = args.c5_learned_chunk
    vh = mat.potential_spectrum(S_f, torch.empty((2048, 2048), dtype=torch.complex64, device=dev))
    out = torch.empty((P, 2048, 2048), dtype=torch.float32, device=dev)

    def timed(fn):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t1, r

    K.run_spectral(corr, mat, pos[:chunk], T, chunk=chunk, vhat=vh, tau=None)      # warm-up
    t_fine, _ = timed(lambda: mat.cbed(pos, S_f, None, out, group=G))
    fine = out.cpu().numpy()
    t_coarse, _ = timed(lambda: mat.cbed(pos, S_f // 8, 8, out, group=G))
    coarse = out.cpu().numpy()
    t_learned, cb = timed(lambda: K.run_spectral(corr, mat, pos, T, chunk=chunk, vhat=vh, tau=None))
    learned = cb.cpu().numpy()
    del cb
    e_c = np.array([np.linalg.norm(coarse[b] - fine[b]) / np.linalg.norm(fine[b]) for b in range(P)])
    e_l = np.array([np.linalg.norm(learned[b] - fine[b]) / np.linalg.norm(fine[b]) for b in range(P)])
    rates = {"fine": P / t_fine, "coarse": P / t_coarse, "learned": P / t_learned}
    log(f"config5 learned: probe-positions/s fine {rates['fine']:.1f}, coarse {rates['coarse']:.1f}, learned "
        f"{rates['learned']:.1f} ({rates['learned'] / rates['fine']:.2f}x fine); CBED rel-L2 vs fine: coarse mean "
        f"{e_c.mean():.4f} median {np.median(e_c):.4f}, learned mean {e_l.mean():.4f} median {np.median(e_l):.4f}")
    # the corrector forward for one chunk, timed with events on its stream
    net = corr.net
    t_fwd, tflops, gbs = corrector_forward_rate(net, chunk, reps=3)
    res = {
        "metric": METRIC, "value": rates["learned"], "unit": "probe-positions/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": t_learned * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 slices, bf16 MFMA corrector",
        "data": "synthetic (seeded polycrystal; corrector trained on other seeds)",
        "config": {"workload": "config 5: seed-5 polycrystal 128x128x1000 A, 2048^2, 300 kV, 30 mrad; learned "
                               "(dz = 8 A, 125 slices + one spectral correction per probe) vs plain fine (dz = 1 A, "
                               f"1000 slices), {P} raster positions",
                   "learned_over_plain_fine": rates["learned"] / rates["fine"],
                   "probe_positions_per_s": rates,
                   "cbed_relL2_vs_fine": {"coarse_mean": float(e_c.mean()), "coarse_median": float(np.median(e_c)),
                                          "learned_mean": float(e_l.mean()), "learned_median": float(np.median(e_l)),
                                          "fraction_improved": float((e_l < e_c).mean())},
                   "fine_vs_fp64_oracle": "3.8e-5 (tests/full_size_config5.py, two positions)",
                   "chunk": chunk, "group": G},
        "roofline": corrector_roofline(net, chunk, t_fwd, tflops, gbs),
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X only: the multislice hot path shards by probe "
              "position across GPUs without any exchange, and this build runs and reports the "
              "single-GPU case (--gpus 1).", file=sys.stderr)
        sys.exit(2)
    import torch

    import namsa_amd._lib as L

    if args.workload == "config1":
        return bench_config1(args)
    if args.workload == "config3":
        return bench_config3(args)
    if args.workload == "config4":
        return bench_config4(args)
    if args.workload == "config5":
        return bench_config5(args)
    if args.workload == "config5-learned":
        return bench_config5_learned(args)
    if args.workload == "potential3d":
        return bench_potential3d(args)
    if args.workload == "records":
        return bench_records(args)
    if args.workload == "stream":
        return bench_stream(args)

    copy_gbs = copy_calibration() if args.copy_calibration else None
    if copy_gbs is not None:
        log(f"copy calibration: 1 GiB read + 1 GiB write, {copy_gbs:.0f} GB/s")

    log("setup: structure, tables, slices, probe, raster")
    sim, dets = setup(args.probes)
    dev = sim.device
    Ny = Nx = 512
    P = int(sim.num_probes)
    S = int(sim.num_slices)
    plane = Ny * Nx
    st = L.stream_handle()
    pos = torch.from_numpy(np.ascontiguousarray(sim.probe_positions, dtype=np.float32)).to(dev)
    cbed = torch.empty((P, Ny, Nx), dtype=torch.float32, device=dev)
    det = torch.from_numpy(dets).to(dev)
    lib = L.load()
    det_ws = torch.empty(int(lib.namsa_det_workspace_doubles(P, Ny, Nx, 2)), dtype=torch.float64, device=dev)
    det_out = torch.empty((P, 2), dtype=torch.float64, device=dev)
    scale = np.float32(1.0 / sim.normalization)
    kmax = np.float32(sim.kmax)

    from namsa_amd import autotune
    from namsa_amd.msa import cache_group

    lo, hi = sim.band
    group = cache_group(Ny, Nx)
    if args.tune:
        log("tuning slice-step schedules at the probe group, then the group")
        tuned = autotune.tune(Nx, batch=min(P, group), reps=5, write_cache=False, band=(lo, hi))
        log(f"tuned: {tuned['best']}  row {tuned['row']}  col {tuned['col']}")
        tg = autotune.tune_group(Nx, P, slices=6, write_cache=False)
        group = int(tg["best"])
        log(f"group times (ms, 6 slices): {tg['times_ms']} -> {group}")
    sched = {"row": autotune.schedules(Nx, 0)[autotune.selected(Nx, 0)],
             "col": autotune.schedules(Nx, 1)[autotune.selected(Nx, 1)], "group": group,
             "band": [lo, hi]}
    ws = torch.empty((min(group, P), Ny, Nx), dtype=torch.complex64, device=dev)

    def step():
        L.call("namsa_multislice", L.ptr(sim.psi_k_device), L.ptr(pos), P, Ny, Nx, kmax, L.ptr(sim._t), S,
               L.ptr(sim.pm_device), lo, hi, group, L.OUT_CBED, scale, L.ptr(ws), L.ptr(cbed), L.ptr(det),
               2, L.ptr(det_ws), L.ptr(det_out), st)

    log(f"warmup {args.warmup} step(s): {P} probes x {S} slices, groups of {group}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} step(s)")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    value = P * S * args.steps / wall
    log(f"wall {wall:.3f} s, events {gpu_ms / 1e3:.3f} s, {value:.4g} probe-slices/s")
    sane = np.isfinite(det_out.cpu().numpy()).all()

    G = min(group, P)
    rf = pass_roofline(sim, pos, ws, G, lo, hi, sched["col"], args.kernel_reps)
    dom = rf["kernel"]
    traffic = None
    tf = ROOT / "profiles" / "hbm_traffic_config2_512.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(dom)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(sim, args.cpu_probes)
        log(f"cpu baseline {cpu['value']:.4g} {cpu['unit']} ({cpu['sample']})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "probe-positions·slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded structure, generated on the fly)",
        "config": {"workload": "batched 4D-STEM plain multislice: 64x64x300 A amorphous+crystal "
                               "(73,728 atoms, Z in {6,14,29,79}), 512x512 complex64, dz=2 A "
                               "(150 slices), 200 kV, 25 mrad, 64x64 raster, CBED + BF/ADF",
                   "grid": [Ny, Nx], "slices": S, "probe_positions": P, "batch": P,
                   "method": "plain", "schedules": sched},
        "roofline": {**rf, "traffic": traffic, "copy_GBs": copy_gbs},
        "cpu_baseline": cpu,
        "finite_outputs": bool(sane),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
