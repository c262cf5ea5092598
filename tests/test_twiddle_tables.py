"""The FFT kernels' twiddle tables, evaluated by the compiler (fft.hpp), against
numpy: for every entry of a few schedules, the stored power of
exp(-2 pi i k / (NS R)) for the thread, butterfly and stage that the Stockham
layout assigns it (t + T q mod NS), rounded to the float pair whose magnitude
is closest to 1.  Entries must equal numpy's cos/sin so rounded, except where
numpy's value is a round-off residue of an exact zero (the tables give 0)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else None)


def _unit_pair(c, s):
    c0, s0 = np.float32(c), np.float32(s)
    c1 = np.nextafter(c0, np.float32(-2 if float(c0) > c else 2))
    s1 = np.nextafter(s0, np.float32(-2 if float(s0) > s else 2))
    best, err = (c0, s0), abs(float(c0) ** 2 + float(s0) ** 2 - 1)
    for ci in (c0, c1):
        for si in (s0, s1):
            e = abs(float(ci) ** 2 + float(si) ** 2 - 1)
            if e < err:
                best, err = (ci, si), e
    return best


def _reference(N, E, radices, compact):
    """Entry (OFF + q P + r - 1) * T + t, stage by stage (no entries for the first)."""
    T = N // E
    out = np.zeros((0, T, 2), np.float32)
    ns = 1
    for R in radices:
        if ns > 1:
            # compact: the powers of two of w up to w^(R-1) for R > 4, else w alone
            P = (4 if R > 8 else 3 if R > 4 else 1) if compact else R - 1
            rows = []
            for q in range(E // R):
                for r in range(1, P + 1):
                    p = (1 << (r - 1)) if compact and R > 4 else r
                    k = (np.arange(T) + T * q) % ns
                    turns = ((-p * k) % (ns * R)) / (ns * R)
                    rows.append([_unit_pair(np.cos(2 * np.pi * f), np.sin(2 * np.pi * f)) for f in turns])
            out = np.concatenate([out, np.array(rows, np.float32)])
        ns *= R
    return out.reshape(-1, 2)


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("tw") / "twiddle_dump"
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'namsa_amd' / 'csrc'}", str(ROOT / "tests" / "twiddle_dump.hip"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    text = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    out = {}
    for block in text.split("# ")[1:]:
        lines = block.strip().splitlines()
        n, e, rs, ct, ntw, t = lines[0].split()
        key = (int(n), int(e), tuple(int(x) for x in rs.split(",")), bool(int(ct)))
        vals = np.array([[float(x) for x in ln.split()] for ln in lines[1:]], np.float32)
        assert vals.shape == (int(ntw) * int(t), 2)
        out[key] = vals
    return out


def test_tables_match_numpy(tables):
    assert len(tables) == 12
    for (N, E, radices, compact), got in tables.items():
        ref = _reference(N, E, radices, compact)
        assert got.shape == ref.shape, (N, radices, compact)
        diff = got != ref
        # only exact zeros may differ from numpy's residues (|cos(pi/2)| ~ 6e-17)
        assert np.all(got[diff] == 0) and np.all(np.abs(ref[diff]) < 1e-15), (N, radices, compact)
        mag = np.hypot(got[:, 0].astype(np.float64), got[:, 1].astype(np.float64))
        assert np.abs(mag - 1).max() < 3e-8, (N, radices, compact)


def test_tables_cover_stage_roots(tables):
    # exact schedule: the last stage's first butterfly row holds w^r for
    # w = exp(-2 pi i k / N), k = t + T q mod N / R, so thread 1 of 512 = 8*8*8
    # has the 512th roots of unity to the powers r * 1
    got = tables[(512, 8, (8, 8, 8), False)]
    T = 64
    # stage 3 (NS = 64) starts after stage 2's 7 entries per thread
    for r in range(1, 8):
        c, s = got[(7 + r - 1) * T + 1]
        assert abs(c - np.cos(2 * np.pi * r / 512)) < 1e-7 and abs(s + np.sin(2 * np.pi * r / 512)) < 1e-7
