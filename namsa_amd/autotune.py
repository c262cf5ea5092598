"""Schedule autotuner for the fused slice step.

Every transform length has several compiled row-pass and column-pass
schedules (radix list, elements per thread, column tile width, next-tile
prefetch, non-temporal traffic; see namsa_amd/csrc/passes_n.hip).  ``tune``
times each schedule on the caller's batch shape with HIP events on the
current stream, keeps the fastest, and records the choice in a JSON cache
keyed by device name and transform length, so later processes start from the
tuned schedule (``apply_cache``).

    from namsa_amd import autotune
    best = autotune.tune(512, batch=4096)     # {'row': name, 'col': name, ...}
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import call, ptr, stream_handle

PASS_ROW, PASS_COL = 0, 1
CACHE_ENV = "NAMSA_TUNE_CACHE"


def cache_path() -> Path:
    p = os.environ.get(CACHE_ENV)
    if p:
        return Path(p)
    return Path.home() / ".cache" / "namsa_amd" / "schedules.json"


def schedules(n: int, pass_: int) -> list[str]:
    lib = _lib.load()
    return [lib.namsa_sched_name(n, pass_, v).decode() for v in range(lib.namsa_sched_count(n, pass_))]


def selected(n: int, pass_: int) -> int:
    lib = _lib.load()
    v = lib.namsa_sched_get(n, pass_)
    if v < 0:
        raise _lib.NamsaError(lib.namsa_last_error().decode())
    return v


def precise(n: int, pass_: int, v=None) -> int:
    """The schedule the fused driver runs instead of v (default: the selected
    one) for thick runs: the form holding every twiddle of a compact one."""
    lib = _lib.load()
    v = selected(n, pass_) if v is None else v
    r = lib.namsa_sched_precise(n, pass_, int(v))
    if r < 0:
        raise _lib.NamsaError(lib.namsa_last_error().decode())
    return r


def driver_schedules(n: int, slices: int) -> tuple:
    """(row, column) schedule indices namsa_multislice runs for `slices` slices."""
    lib = _lib.load()
    thick = slices > lib.namsa_precise_slices(n)
    return tuple(precise(n, p) if thick else selected(n, p) for p in (PASS_ROW, PASS_COL))


def select(n: int, pass_: int, v) -> None:
    if isinstance(v, str):
        v = schedules(n, pass_).index(v)
    call("namsa_sched_set", int(n), int(pass_), int(v))


def _device_key() -> str:
    import torch

    return torch.cuda.get_device_name(torch.cuda.current_device())


def _time(fn, reps: int, reset=None) -> float:
    """Median over 3 trials of the mean time of `reps` back-to-back calls (ms),
    bracketed by one pair of events (no gaps between the launches).  `reset`
    (untimed) runs before the warm-up and each trial."""
    import torch

    if reset is not None:
        reset()
    fn()
    fn()
    trials = []
    for _ in range(3):
        if reset is not None:
            reset()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        trials.append(e0.elapsed_time(e1) / reps)
    return float(np.median(trials))


def tune(n: int, batch: int = 0, reps: int = 20, write_cache: bool = True, verbose: bool = False,
         band=None, compact: bool = True) -> dict:
    """Select the fastest schedule of each pass at n x n with `batch` waves.

    Passes are timed the way the slice loop runs them: `reps` slice steps
    (row pass, then column pass) back to back.  Column variants are swept
    with the current row variant, then row variants with the best column one
    (the two passes share no state but the wave).  Returned times are per
    slice step.

    batch: waves per launch (default: the cache-sized probe group production uses).
    band: the propagator's open column range (default: the 1/3-bandwidth band).
    compact: also consider the compact-twiddle (.ct) schedules (fewer
    registers, 9-27% faster; their round-off grows 1.2-1.4x as fast with the
    slice count as that of the schedules holding every twiddle).
    """
    import torch

    if batch <= 0:
        from .msa import cache_group

        batch = cache_group(n, n)

    dev = torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(0)
    # Every variant starts from the same random wave: the slice step runs
    # about 20% faster on zeros than on non-zero values, and a wave pushed
    # through many random slices decays towards zeros, which would favour the
    # variants timed last.
    psi0 = torch.randn((batch, n, n), dtype=torch.complex64, device=dev)
    psi = psi0.clone()
    ph = torch.from_numpy(rng.uniform(0, 2 * np.pi, (4, n, n)).astype(np.float32)).to(dev)
    t = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    pm = t[0].flip(0).contiguous()
    st = stream_handle()
    lo, hi = band if band is not None else default_band(n)
    # the wave layout namsa_multislice's slice loop uses at this size
    sfx = "_tiled" if _lib.load().namsa_wave_tiled(n, n) else ""
    result = {"n": n, "batch": batch, "band": [lo, hi], "row": {}, "col": {}, "tiled": bool(sfx)}
    i = [0]

    def step():
        call("namsa_slice_rows" + sfx, ptr(psi), batch, n, n, ptr(t[1 + i[0] % 3]), lo, hi, st)
        call("namsa_slice_cols" + sfx, ptr(psi), batch, n, n, ptr(pm), lo, hi, st)
        i[0] += 1

    for pass_, key in ((PASS_COL, "col"), (PASS_ROW, "row")):
        for v, name in enumerate(schedules(n, pass_)):
            if name.endswith(".ct") and not compact:
                continue
            select(n, pass_, v)
            result[key][name] = _time(step, reps, reset=lambda: psi.copy_(psi0))
            if verbose:
                print(f"  n={n} {key} {name}: {result[key][name]:.4f} ms per slice step", flush=True)
        best = min(result[key], key=result[key].get)
        select(n, pass_, best)
    result["best"] = {"row": min(result["row"], key=result["row"].get),
                      "col": min(result["col"], key=result["col"].get)}
    if write_cache:
        _store(n, result["best"])
    return result


def default_band(n: int) -> tuple:
    """Open columns of the default 1/3-bandwidth mask on an n-point linspace grid."""
    k = np.arange(n) / (n - 1) - 0.5                 # in units of kmax
    kmin_y = np.min(np.abs(k))
    open_cols = np.nonzero(np.sqrt(k * k + kmin_y * kmin_y) < 1 / 3)[0]
    return int(open_cols[0]), int(open_cols[-1]) + 1


def tune_group(n: int, probes: int, slices: int = 8, candidates=None, write_cache: bool = True,
               verbose: bool = False) -> dict:
    """Time the fused multislice at several probe-group sizes; select the fastest.

    Groups whose waves fit the last-level cache keep the slice loop out of HBM;
    the best size depends on the cache, the grid and the schedules.
    """
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    plane = n * n
    base = max(1, (256 << 20) // (8 * plane))
    if candidates is None:
        candidates = sorted({max(1, int(base * f)) for f in (0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0, 1.5)}
                            | {probes})
    candidates = [c for c in candidates if c <= probes]
    rng = np.random.default_rng(1)
    ph = torch.from_numpy(rng.uniform(0, 2 * np.pi, (slices, n, n)).astype(np.float32)).to(dev)
    t = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    lo, hi = default_band(n)
    pm = torch.zeros((n, n), dtype=torch.complex64, device=dev)
    pm[:, lo:hi] = t[0, :, lo:hi]
    psi_k = t[-1].clone()
    pos = torch.zeros((probes, 2), dtype=torch.float32, device=dev)
    cbed = torch.empty((probes, n, n), dtype=torch.float32, device=dev)
    ws = torch.empty((max(candidates), n, n), dtype=torch.complex64, device=dev)
    st = stream_handle()
    times = {}
    for g in candidates:
        fn = lambda: call("namsa_multislice", ptr(psi_k), ptr(pos), probes, n, n, np.float32(6.0), ptr(t),
                          slices, ptr(pm), lo, hi, g, _lib.OUT_CBED, np.float32(1.0), ptr(ws), ptr(cbed),
                          None, 0, None, None, st)
        times[g] = _time(fn, 2)
        if verbose:
            print(f"  n={n} group={g}: {times[g]:.2f} ms", flush=True)
    best = min(times, key=times.get)
    if write_cache:
        _store(n, {"group": int(best)}, merge=True)
    return {"n": n, "probes": probes, "slices": slices, "times_ms": times, "best": best}


def cached_group(Ny: int, Nx: int):
    """The tuned probe-group size for this device and grid, or None."""
    if Ny != Nx:
        return None
    try:
        data = json.loads(cache_path().read_text())
        return int(data.get(_device_key(), {}).get(str(Nx), {}).get("group", 0)) or None
    except (OSError, ValueError, RuntimeError):
        return None


def _store(n: int, best: dict, merge: bool = False) -> None:
    path = cache_path()
    try:
        data = json.loads(path.read_text()) if path.exists() else {}
    except (OSError, ValueError):
        data = {}
    entry = data.setdefault(_device_key(), {})
    if merge:
        entry.setdefault(str(n), {}).update(best)
    else:
        entry[str(n)] = {**entry.get(str(n), {}), **best}
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache, never the selection


def apply_cache() -> dict:
    """Select the cached schedules for this device; returns what was applied."""
    path = cache_path()
    if not path.exists():
        return {}
    try:
        data = json.loads(path.read_text()).get(_device_key(), {})
    except (OSError, ValueError):
        return {}
    applied = {}
    for n, best in data.items():
        try:
            if "row" in best:
                select(int(n), PASS_ROW, best["row"])
            if "col" in best:
                select(int(n), PASS_COL, best["col"])
            applied[int(n)] = best
        except (ValueError, KeyError, _lib.NamsaError):
            continue  # schedule names from another build
    return applied
