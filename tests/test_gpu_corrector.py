"""Coarse-slice corrector on MI355X: the BF16 MFMA network against the FP64
oracle (forward and every gradient group), the coarse/fine stepping and
readout against the plain path, the fallback, and the gain of a trained
corrector on held-out structures.

Tolerances: BF16 operands (8-bit mantissa) with FP32 accumulation through
four Fourier layers: forward output within 2e-2 relative L2 of the FP64
oracle, each gradient group within 5e-2."""
import sys
import time

import numpy as np
import pytest

from namsa_amd import corrector as K
from namsa_amd import structures
from namsa_amd.msa import MSAGPU
from oracle import fno_oracle as F

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("cfg", [
    dict(H=32, W=32, cin=3, C=8, L=2, m=8, Hq=16, B=2),
    dict(H=64, W=64, cin=10, C=32, L=4, m=16, Hq=64, B=2),   # production widths and modes
])
def test_network_matches_oracle(gpu, cfg):
    import torch
    cfg = dict(cfg)
    B = cfg.pop("B")
    _say(f"corrector network vs oracle {cfg} B={B}")
    rng = np.random.default_rng(11)
    net = K.FnoNet(**cfg, seed=3)
    p = net.params_numpy()
    for k in ("bl", "b", "bq1", "bq2"):
        p[k] = rng.standard_normal(p[k].shape) * 0.2
    p["Q2"] = rng.standard_normal(p["Q2"].shape) / np.sqrt(cfg["Hq"])   # the default init zeroes it
    net.load(p)
    p = net.params_numpy()
    H, W, cin = cfg["H"], cfg["W"], cfg["cin"]
    u = rng.standard_normal((B, cin, H, W))
    gout = rng.standard_normal((B, 2, H, W))
    net.ensure(B)
    net.u[:, :B * H * W].copy_(torch.from_numpy(u.transpose(1, 0, 2, 3).reshape(cin, -1).astype(np.float32)))
    net.forward(B)
    net.gout[:, :B * H * W].copy_(torch.from_numpy(gout.transpose(1, 0, 2, 3).reshape(2, -1).astype(np.float32)))
    net.backward(B)
    torch.cuda.synchronize()
    out = net.out[:, :B * H * W].cpu().numpy().reshape(2, B, H, W).transpose(1, 0, 2, 3)
    ref, cache = F.forward(p, u)
    e = _rel(out, ref)
    _say(f"  forward relL2 {e:.2e}")
    assert e < 2e-2
    g = net.grads_numpy()
    gref = F.backward(p, cache, gout)
    errs = {k: _rel(g[k], gref[k]) for k in net.shapes}
    _say("  grads relL2 " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
    assert max(errs.values()) < 5e-2


def _material(seed, thickness, factor=8):
    rec = structures.random_amorphous_crystal(seed=seed, lateral=32.0, thickness=thickness)
    sim = MSAGPU(200, 0.025, rec, sampling=np.array([256, 256]), debye_waller=False)
    sim.setup_device(0)
    sim.calc_atomic_potentials()
    return K.CoarseFineMaterial(sim, dz=1.0, factor=factor)


def _positions(seed, n):
    import torch
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.uniform(-8, 8, size=(n, 2)).astype(np.float32)).cuda()


def test_stepwise_readout_matches_fused_path(gpu):
    """Coarse steps through the per-slice entry points, the real-space
    transform and the masked intensity readout reproduce the fused driver's
    CBEDs, which MSAGPU itself returns for the same slices."""
    import torch
    _say("stepwise coarse stepping + real-space readout vs fused driver")
    mat = _material(101, 48.0)
    pos = _positions(1, 3)
    ws = torch.empty((3, mat.Ny, mat.Nx), dtype=torch.complex64, device="cuda")
    mat.init(ws, pos)
    for k in range(mat.S_c):
        mat.step(ws, mat.t_c[k], mat.pm_c)
    mat.to_real(ws)
    cb = torch.empty((3, mat.Ny, mat.Nx), dtype=torch.float32, device="cuda")
    mat.cbed_from_real(ws, cb)
    fused = K.run_plain(mat, pos, coarse=True).cpu().numpy()
    cb = cb.cpu().numpy()
    sim = mat.sim
    sim.probe_positions = pos.cpu().numpy().astype(np.float64)
    sim.num_probes = 3
    sim.plan_simulation()
    sim.multislice()
    e = max(_rel(cb[i], fused[i]) for i in range(3))
    e2 = max(_rel(fused[i], sim.probes[i]) for i in range(3))
    _say(f"  S_c={mat.S_c} stepwise vs fused relL2 {e:.2e}, fused vs MSAGPU {e2:.2e}")
    assert e < 1e-5 and e2 < 1e-6


def test_trained_corrector_gain_and_fallback(gpu):
    """Training lowers the per-step residual on held-out structures (the gain
    the network is trained for); the end-to-end CBED error of the learned
    path is measured and reported beside the plain coarse path's, and the
    fallback with tau = 0 reproduces the fine path exactly."""
    import torch
    _say("corrector training on coarse/fine pairs, 256^2")
    t0 = time.time()
    rng = np.random.default_rng(0)
    pool = K.PairPool()
    for i in range(6):
        pool.add_material(_material(200 + i, float(8 * rng.integers(5, 12))), _positions(10 + i, 8))
    pool.finalize()
    held = K.PairPool()
    hmat = _material(300, 64.0)
    hpos = _positions(99, 4)
    held.add_material(hmat, hpos)
    held.finalize()
    torch.cuda.synchronize()
    _say(f"  pairs: train {len(pool)} held-out {len(held)} ({time.time() - t0:.1f} s)")
    net = K.FnoNet(256, 256, cin=hmat.n_inputs(True), C=32, L=4, m=16, Hq=64, seed=0, derivs=True)
    base, _ = K.evaluate(net, held, batch=8)
    assert abs(base - 1.0) < 1e-6          # zero output layer: no correction yet
    t0 = time.time()
    K.train(net, pool, steps=1000, batch=8, lr=2e-3, seed=0, log_every=250, log=_say)
    torch.cuda.synchronize()
    _say(f"  training 1000 steps: {time.time() - t0:.1f} s")
    tr, _ = K.evaluate(net, pool, batch=8)
    ho, per = K.evaluate(net, held, batch=8)
    _say(f"  relative residual: train {tr:.3f}  held-out {ho:.3f} (max per pair {per.max():.3f}; "
         f"1.0 = coarse step uncorrected)")
    assert ho < 0.9
    fine = K.run_plain(hmat, hpos).cpu().numpy()
    coarse = K.run_plain(hmat, hpos, coarse=True).cpu().numpy()
    learned, nf = K.run_learned(net, hmat, hpos)
    assert nf == 0
    learned = learned.cpu().numpy()
    ec = max(_rel(coarse[i], fine[i]) for i in range(4))
    el = max(_rel(learned[i], fine[i]) for i in range(4))
    _say(f"  CBED relL2 vs fine (reported): coarse {ec:.3e}  learned {el:.3e}")
    assert np.all(np.isfinite(learned))
    fb, nf = K.run_learned(net, hmat, hpos, tau=0.0)
    assert nf == hmat.S_c
    e = max(_rel(fb.cpu().numpy()[i], fine[i]) for i in range(4))
    _say(f"  fallback everywhere: {nf} slices, relL2 vs fine {e:.1e}")
    assert e < 1e-6
