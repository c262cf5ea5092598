"""CPU baseline: op-for-op port of the reference CPU ``MSA`` numpy path -- TEST INFRASTRUCTURE.

Used only as the timed ``cpu_baseline`` leg of bench.py and by tests; never by
the product package.  It follows namsa/msa.py:320-361 (``propagate_beam``,
numpy branch, ``probe_grid=True``) with its helpers:

  build_propagator      msa.py:238-243   exp(-i pi lambda dz k^2), complex64
  bandwidth_limit_mask  msa.py:245-251   integer-pixel disc, float32
  build_probe (CPU)     msa.py:198-236   ifft2(norm='ortho') + fftshift, |psi|=1
  slices                msa.py:334       exp(1j*sigma*V) of ALL slices, per probe
  slice loop            msa.py:346-353   descending order, t*psi -> fft2 ->
                                         *mask*prop -> ifft2
  multislice            msa.py:289-312   multiprocessing Pool over probes,
                                         processes = min(cpu_count, n_probes)

Arrays stay complex64 as in the reference (numpy >= 2 keeps complex64 through
its FFT; the reference era's numpy 1.x upcast to complex128).

Pinning: tests/test_oracle.py checks this port against msa_oracle in the
regime where the CPU and GPU conventions coincide (probe at the origin, full
bandwidth mask, slice stack symmetric under reversal, equal aperture
steepness), so the arithmetic is pinned; its CPU-only quirks (integer-pixel
mask, (x, y) position order, descending order) are pinned by the
closed-form checks there.
"""
from __future__ import annotations

import multiprocessing as mp
import os

import numpy as np

_SHARED = {}


def _k_mesh(kmax, Ny, Nx):
    ky = np.linspace(-kmax / 2, kmax / 2, Ny)
    kx = np.linspace(-kmax / 2, kmax / 2, Nx)
    return np.meshgrid(ky, kx, indexing="ij")


def build_propagator(kmax, Ny, Nx, lam, dz):
    k_y, k_x = _k_mesh(kmax, Ny, Nx)
    return np.exp(-np.pi * 1j * lam * dz * (k_x ** 2 + k_y ** 2)).astype(np.complex64)


def bandwidth_limit_mask(shape, radius):
    g0, g1 = np.meshgrid(np.arange(-shape[0] // 2, shape[0] // 2),
                         np.arange(-shape[1] // 2, shape[1] // 2), indexing="ij")
    r = np.sqrt(g0 ** 2 + g1 ** 2)
    return np.heaviside(max(shape) * radius - r, 0).astype(np.float32)


def build_probe_cpu(kmax, Ny, Nx, lam, semi_angle, pos, smooth_apert=True,
                    apert_smooth=50.0, C1=0.0, C3=0.0, C5=0.0, scherzer=True):
    k_y, k_x = _k_mesh(kmax, Ny, Nx)
    k_rad = np.sqrt(k_x ** 2 + k_y ** 2)
    k_semi = semi_angle / lam
    if smooth_apert:
        # exp overflows to inf far outside the aperture, where 1 / (1 + inf) = 0 is the value wanted
        with np.errstate(over="ignore"):
            apert = 1 / (1 + np.exp(-2 * apert_smooth * (k_semi - k_rad)))
    else:
        apert = np.heaviside(k_semi - k_rad, 0.5)
    if scherzer and C3 > 1.0:
        C1 = np.sqrt(1.5) * np.sqrt(C3 * lam)
    kl = k_rad * lam
    chi = 2 * np.pi / lam * (-0.5 * C1 * kl ** 2 + 0.25 * C3 * kl ** 4 + C5 / 6 * kl ** 6)
    psi_k = (apert * np.exp(-1j * chi)).astype(np.complex64)
    x, y = pos                                     # msa.py:225 takes (x, y)
    ramp = np.exp(2j * np.pi * (-k_x * x - k_y * y)).astype(np.complex64)
    psi = np.fft.fftshift(np.fft.ifft2(psi_k * ramp, norm="ortho"))
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2))
    return psi.astype(np.complex64)


def cbed_readout(psi):
    """The CBED the reference's notebooks take of a CPU exit wave:
    |fftshift(fft2(psi))|^2, unnormalised, float32."""
    return (np.abs(np.fft.fftshift(np.fft.fft2(psi))) ** 2).astype(np.float32)


def propagate_beam(pos, V, sigma, kmax, lam, dz, semi_angle, bandwidth, probe_kw, cbed=False):
    """One probe through all slices (msa.py:320-361, numpy branch); with
    cbed=True also the readout of the exit wave (cbed_readout)."""
    Ny, Nx = V.shape[1:]
    propag = build_propagator(kmax, Ny, Nx, lam, dz)
    blim = bandwidth_limit_mask((Ny, Nx), bandwidth)
    psi = build_probe_cpu(kmax, Ny, Nx, lam, semi_angle, pos, **probe_kw)
    slices = np.exp(1j * sigma * V).astype(np.complex64)
    for trans in slices[::-1]:
        psi = np.fft.ifft2(np.fft.fft2(trans * psi) * blim * propag)
    return cbed_readout(psi) if cbed else psi


def _worker(args):
    return propagate_beam(args, _SHARED["V"], *_SHARED["rest"])


def multislice_cpu(positions, V, sigma, kmax, lam, dz, semi_angle,
                   bandwidth=1 / 3, probe_kw=None, processes=None, cbed=False):
    """Exit waves (or, with cbed=True, CBEDs) for all positions with a fork
    Pool (msa.py:289-312)."""
    probe_kw = probe_kw or {}
    positions = [tuple(p) for p in np.asarray(positions, dtype=np.float64)]
    if processes is None:
        processes = min(os.cpu_count() or 1, len(positions))
    if processes <= 1:
        return np.array([propagate_beam(p, V, sigma, kmax, lam, dz, semi_angle,
                                        bandwidth, probe_kw, cbed) for p in positions])
    _SHARED["V"] = np.asarray(V, dtype=np.float32)
    _SHARED["rest"] = (sigma, kmax, lam, dz, semi_angle, bandwidth, probe_kw, cbed)
    chunk = max(1, len(positions) // processes)
    ctx = mp.get_context("fork")
    with ctx.Pool(processes=processes) as pool:
        out = pool.map(_worker, positions, chunksize=chunk)
    _SHARED.clear()
    return np.array(out)
