"""Train the spectral coarse-slice corrector on pairs the plain method generates.

    python -m namsa_amd.train_corrector --train 2250 --held-out 250 --factors 2,4,8 \
        --steps 6000 --out namsa_amd/data --report report.json

Training pairs (the configuration the corrector is specified on): seeded
random crystals (cubic / tetragonal / hexagonal lattices, a in [3, 7] A, 1-8
basis atoms, Z in [3, 83]) 32 A wide at a 256 x 256 grid, 200 kV, 25 mrad,
40-400 A thick; coarse multislice at dz = f A (f = 8 by default) against
fine multislice at dz = 1 A.  Each structure gives 4 probe positions at two
thicknesses, so --train 2250 --held-out 250 is the 18k / 2k pair budget.
Held-out structures come from a disjoint seed range.  One network is trained
per coarsening factor on the same fine targets; each is saved as
corrector_<tag>_f<f>.npz (bfloat16 weights) and scored on the held-out pairs.

--preset poly: the large-scale comparison's structure type instead -- 4-grain
Si / GaAs / Au polycrystals 128 x 128 x 1000 A at 2048^2, 300 kV, 30 mrad,
snapshots at 1000 A and a random multiple of 8 A from 400 A up, seeds from
200 on (the benchmark's structure is seed 5).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

TRAIN_SEED0 = 100_000
HELD_SEED0 = 900_000


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--train", type=int, default=2250, help="training structures (8 pairs each)")
    p.add_argument("--held-out", type=int, default=250, help="held-out structures (8 pairs each)")
    p.add_argument("--factors", default="8", help="coarsening factors, comma separated")
    p.add_argument("--steps", type=int, default=6000)
    p.add_argument("--batch", type=int, default=16)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--grid", type=int, default=256)
    p.add_argument("--opts", type=int, default=1, help="input options: 1 potential spectrum, 2 amplitude")
    p.add_argument("--tag", default=None)
    p.add_argument("--preset", choices=["crystal", "poly"], default="crystal")
    p.add_argument("--probes", type=int, default=None, help="positions per structure")
    p.add_argument("--out", default=None, help="directory for the weight files (none: not saved)")
    p.add_argument("--report", default=None, help="JSON report path")
    args = p.parse_args(argv)

    import torch

    from namsa_amd.corrector import SpectralCorrector, make_spectral_pools, spectral_evaluate, spectral_train

    from namsa_amd.corrector import polycrystal_sim

    factors = tuple(int(f) for f in args.factors.split(","))
    t0 = time.perf_counter()
    log(f"pairs: {args.train} training structures ({args.preset}), factors {factors}")
    if args.preset == "poly":
        tag = args.tag or "poly"
        probes = args.probes or 32
        kw = dict(probes=probes, thickness=(400.0, 1000.0), builder=lambda sd: polycrystal_sim(sd))
        seeds, held_seeds = range(200, 200 + args.train), range(900, 900 + args.held_out)
    else:
        tag = args.tag or "c4"
        probes = args.probes or 4
        kw = dict(probes=probes, grid=args.grid)
        seeds = range(TRAIN_SEED0, TRAIN_SEED0 + args.train)
        held_seeds = range(HELD_SEED0, HELD_SEED0 + args.held_out)
    pool, mat = make_spectral_pools(seeds, factors=factors, log=log, **kw)
    t_pool = time.perf_counter() - t0
    log(f"  {len(pool)} training pairs in {t_pool:.0f} s")
    held, _ = make_spectral_pools(held_seeds, factors=factors, log=log, **kw)
    log(f"  {len(held)} held-out pairs")
    grid = mat.Ny
    report = {"preset": args.preset, "train_pairs": len(pool), "held_out_pairs": len(held), "pool_seconds": t_pool,
              "factors": {}, "grid": grid, "steps": args.steps, "batch": args.batch, "opts": args.opts}
    for f in factors:
        corr = SpectralCorrector(grid, grid, mat.window, mat.sim.kmax, mat.lam, f * mat.dz, opts=args.opts,
                                 factor=f, device=mat.sim.device)
        log(f"factor {f}: window {mat.window}, {corr.net.n_params} parameters, {args.steps} steps of {args.batch}")
        t1 = time.perf_counter()
        hist = spectral_train(corr, pool, args.steps, batch=args.batch, lr=args.lr, seed=f,
                              log_every=max(1, args.steps // 20), log=log)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        learned, coarse = spectral_evaluate(corr, held)
        r = {"train_seconds": t_train, "held_out_learned_mean": float(learned.mean()),
             "held_out_learned_median": float(np.median(learned)), "held_out_coarse_mean": float(coarse.mean()),
             "held_out_coarse_median": float(np.median(coarse)), "fraction_improved": float((learned < coarse).mean()),
             "history": hist}
        log(f"factor {f}: held-out CBED rel-L2 learned mean {r['held_out_learned_mean']:.4f} median "
            f"{r['held_out_learned_median']:.4f} | coarse mean {r['held_out_coarse_mean']:.4f} median "
            f"{r['held_out_coarse_median']:.4f} | improved {r['fraction_improved']:.2f} ({t_train:.0f} s)")
        if args.out:
            path = f"{args.out}/corrector_{tag}_f{f}.npz"
            corr.save(path)
            r["weights"] = path
        report["factors"][str(f)] = r
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)
    return report


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
