"""Coarse-slice corrector network on MI355X: the BF16 MFMA forward and every
gradient group against the FP64 oracle (the spectral corrector built on it
is tested in test_gpu_spectral.py).

Tolerances: BF16 operands (8-bit mantissa) with FP32 accumulation through
four Fourier layers: forward output within 2e-2 relative L2 of the FP64
oracle, each gradient group within 5e-2."""
import sys

import numpy as np
import pytest

from namsa_amd import corrector as K
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


@pytest.mark.parametrize("H,W,B", [
    (37, 70, 3),      # 111 rows: the last workgroup has a masked wavefront; W = 70 ends in a partial tile
    (64, 64, 2),
    (170, 170, 2),    # the config-4 window
])
def test_fused_forward_matches_oracle(gpu, H, W, B):
    """Channel-last BF16 inference forward against the FP64 oracle and the
    FP32-activation forward of the same weights.  Tolerance 3e-2: BF16
    operands as in the strided path, plus activations rounded to bf16 between
    layers."""
    import torch
    _say(f"fused forward vs oracle H={H} W={W} B={B}")
    rng = np.random.default_rng(H + W)
    net = K.FnoNet(H, W, cin=7, C=32, L=4, m=16, Hq=64, seed=5)
    p = net.params_numpy()
    for k in ("bl", "b", "bq1", "bq2"):
        p[k] = rng.standard_normal(p[k].shape) * 0.2
    p["Q2"] = rng.standard_normal(p["Q2"].shape) / np.sqrt(64)
    net.load(p)
    p = net.params_numpy()
    u = rng.standard_normal((B, 7, H, W))
    n = B * H * W
    net.ensure(B)
    net.u[:, :n].copy_(torch.from_numpy(u.transpose(1, 0, 2, 3).reshape(7, -1).astype(np.float32)))
    fused = net.forward_fused(B)[:, :n].cpu().numpy().reshape(2, B, H, W).transpose(1, 0, 2, 3)
    fp32 = net.forward(B)[:, :n].cpu().numpy().reshape(2, B, H, W).transpose(1, 0, 2, 3)
    ref, _ = F.forward(p, u)
    e_ref, e_fp32 = _rel(fused, ref), _rel(fused, fp32)
    _say(f"  fused vs FP64 oracle {e_ref:.2e}, vs FP32-activation forward {e_fp32:.2e}, FP32 vs oracle "
         f"{_rel(fp32, ref):.2e}")
    assert np.isfinite(fused).all()
    assert e_ref < 3e-2 and e_fp32 < 3e-2


def test_forward_graph_replays_match_eager(gpu):
    """The captured fused forward: every replay equals the eager pass bit for
    bit, with eager work and synchronisations between replays, new inputs
    between replays, and a larger batch (which reallocates) invalidating and
    recapturing the graph."""
    import torch
    rng = np.random.default_rng(8)
    H, W = 48, 80
    net = K.FnoNet(H, W, cin=7, C=32, L=4, m=16, Hq=64, seed=9)
    p = net.params_numpy()
    p["Q2"] = rng.standard_normal(p["Q2"].shape) / 8
    net.load(p)
    for B in (3, 3, 5):
        n = B * H * W
        for rep in range(3):
            net._ensure_common(B)
            u = torch.from_numpy(rng.standard_normal((7, n)).astype(np.float32)).cuda()
            net.u[:, :n].copy_(u)
            eager = net.forward_fused(B)[:, :n].clone()
            _ = float(eager.sum())                      # eager work and a synchronisation
            net.out.fill_(np.nan)
            got = net.forward_graph(B)[:, :n]
            torch.cuda.synchronize()
            assert torch.equal(got, eager), (B, rep)
        _say(f"  B={B}: 3 replays equal to the eager pass")
