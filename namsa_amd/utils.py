"""Helpers with counterparts in the reference's namsa/utils.py, run on the GPU."""
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
