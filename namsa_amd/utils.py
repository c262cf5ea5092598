"""Helpers with counterparts in the reference's namsa/utils.py: the
thin-object approximation (on the GPU) and the host-side crystallography and
array helpers the reference's scripts use."""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import call, ptr, stream_handle


def thin_object_approx(potential_slices, probes, device=None):
    """Thin-object exit amplitudes |FFT2(probes * exp(i sum_s V_s))|
    (namsa/utils.py:130-135), unshifted and unnormalised, through
    namsa_thin_object (phase summed in FP64, transform in FP32).

    potential_slices: [S][Ny][Nx] real; probes: [B][Ny][Nx] or [Ny][Nx] complex.
    Returns float32 of the probes' shape.
    """
    import torch

    V = np.asarray(potential_slices, dtype=np.float64)
    if V.ndim == 2:
        V = V[None]
    P = np.asarray(probes, dtype=np.complex64)
    single = P.ndim == 2
    if single:
        P = P[None]
    if V.shape[1:] != P.shape[1:]:
        raise ValueError(f"slices {V.shape[1:]} and probes {P.shape[1:]} differ in shape")
    _lib.load()
    dev = torch.device("cuda", 0) if device is None else torch.device(device)
    B, Ny, Nx = P.shape
    V_d = torch.from_numpy(np.ascontiguousarray(V)).to(dev)
    P_d = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    ws = torch.empty((B, Ny, Nx), dtype=torch.complex64, device=dev)
    out = torch.empty((B, Ny, Nx), dtype=torch.float32, device=dev)
    call("namsa_thin_object", ptr(V_d), V.shape[0], ptr(P_d), B, Ny, Nx, ptr(ws), ptr(out), stream_handle())
    res = out.cpu().numpy()
    return res[0] if single else res


def dhkl_spacing(uc_volume, basis, hkl=(0, 0, 1)):
    """Interplanar spacing 1 / |h b1 + k b2 + l b3| (Angstrom) from the lattice
    vectors a1, a2, a3 (rows of `basis`) and the cell volume
    (namsa/utils.py:91-98)."""
    a1, a2, a3 = np.asarray(basis, dtype=np.float64)
    recip = np.array([np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)]) / uc_volume
    g = hkl[0] * recip[0] + hkl[1] * recip[1] + hkl[2] * recip[2]
    return 1.0 / np.sqrt(np.dot(g, g))


def bin_2d_array(arr, dwnspl=2, mode="mean"):
    """Block mean (or, with any other mode, block sum) over dwnspl x dwnspl
    blocks (namsa/utils.py:101-108); trailing rows/columns must divide."""
    a = np.asarray(arr)
    r, c = a.shape[0] // dwnspl, a.shape[1] // dwnspl
    blocks = a.reshape(r, dwnspl, c, dwnspl)
    return blocks.mean(-1).mean(-2) if mode == "mean" else blocks.sum(-1).sum(-2)


def imageTile(data, padsize=1, padval=0):
    """Tile [n][h][w...] images into one square mosaic, each image followed by
    padsize rows and columns of padval (namsa/utils.py:111-127)."""
    data = np.asarray(data)
    n = int(np.ceil(np.sqrt(data.shape[0])))
    pad = ((0, n * n - data.shape[0]), (0, padsize), (0, padsize)) + ((0, 0),) * (data.ndim - 3)
    data = np.pad(data, pad, mode="constant", constant_values=(padval, padval))
    h, w = data.shape[1], data.shape[2]
    grid = data.reshape((n, n) + data.shape[1:])
    grid = np.swapaxes(grid, 1, 2)                       # (row of tiles, h, column of tiles, w, ...)
    return grid.reshape((n * h, n * w) + data.shape[3:])
