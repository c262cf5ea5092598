"""FP64 restatement of the reference MSAGPU multislice path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module.  Wave arithmetic is complex128 (numpy pocketfft); discrete decisions
that the reference takes in FP32 (bandwidth mask, hard aperture, slice
membership, site pixel placement) are replicated bit-exactly, because a
flipped mask pixel would break the intensity tolerance on its own.

Conventions (reference file:line each function follows):
  * k-grid is a linspace, k_j = j*kmax/(N-1) - kmax/2 in double from the
    float32 kmax the kernels receive (probe_kernels.cu:11-16).
  * probe = fftshift(IFFT_unnormalised(psi_k * ramp)) (msa.py:959-964,
    probe_kernels.cu:239-253); the ramp phase is taken in FP64 here (the
    reference forms it in FP32).
  * slice step psi <- IFFT_u(PM * FFT(psi * t_i / N^2)), ascending i
    (msa.py:971-981).
  * readout CBED = |sqrt(N^2) FFT(psi / N^2)|^2 / normalization
    (msa.py:985-990, :905).
  * padded sites are skipped (the reference reads atom_pot_stack[-1] for
    them, potential_kernels.cu:25-41; that out-of-bounds read is not
    reproducible and is deliberately not followed).

Pinning (tests/test_oracle.py):
  voltage2lambda, sigma_int ........ textbook KATs at 100/200/300 kV
  projected_potential .............. K0 special value, plane integral
                                      2*pi*a0*e*(sum a/b + sum c) at eps=0,
                                      Gaussian term alone vs its closed form
  atomic_table ..................... shape rule, min == 0, symmetry, 2x2 mean
                                      vs direct subsample average
  bin_sites / splat_slices ......... brute force one-atom placement, edge
                                      clipping, linearity, sentinel skip
  transmission ..................... |t| = 1, NaN -> 1 special case
  geometry ......................... config KATs (64 A at 512^2: kmax 8.0,
                                      pixel 0.125, max angle 200.6 mrad;
                                      38 A at 256^2: 6.74, 0.148) and a
                                      hand-computed anisotropic (py, px) case
  aperture ......................... A = 1/2 exactly at float32 k = k_semi,
                                      1/4 and 3/4 at ln(3)/160 either side
                                      (steepness 2*80), monotone decay
  chi_phase ........................ C1-, C3-, C5-only and combined closed
                                      forms at hand-computed pixel radii
  k_axis / probe_k ................. closed-form spot values, |psi_k| = A
  probe_real / probe_stack ......... Parseval (norm = N^2 sum |psi_k|^2),
                                      position-0 stack == build_probe psi,
                                      brute-force DFT at N = 8
  propagator_mask .................. composition P(a)P(b) = P(a+b) on the mask
  multislice ....................... vacuum closed form mask*|psi_k|^2/sum,
                                      unitarity with a full mask, tilt t ->
                                      shifted CBED, brute-force DFT at N = 8
  probe_positions .................. notebook KAT 17x17 = 289, explicit grid
  detector_mask_reference .......... brute-force annulus from pixel radii
  integrate_slices ................. constant and linear-in-z closed forms,
                                      the intervals between chunks left out,
                                      z crop rule on a 13-sample grid
  integration_params ............... worked example of msa.py:409-421
  thin_object ...................... vacuum = |FFT(probe)|, brute-force DFT
                                      at N = 8
"""
from __future__ import annotations

import math

import numpy as np
from scipy.constants import elementary_charge, electron_mass, speed_of_light
from scipy.special import expit, k0

A0_BOHR = 0.529       # Angstrom, msa.py:183
E_CHARGE = 14.4       # V*Angstrom, msa.py:184
BESSEL_EPS = 0.1      # Angstrom, msa.py:182
PI_F32 = float(np.float32(math.pi))   # acos(-1.f) in spherical_phase_error, probe_kernels.cu:210


# ---------------------------------------------------------------- optics
def voltage2lambda(volts):
    """Relativistic electron wavelength in Angstrom (optics.py:4-12)."""
    return 12.2643 / np.sqrt(volts + 0.97845e-6 * volts ** 2)


def sigma_int(volts):
    """Interaction parameter in rad/(V*Angstrom) (optics.py:100-108)."""
    mc2 = electron_mass * speed_of_light ** 2
    eV = elementary_charge * volts
    return 2 * np.pi * (mc2 + eV) / (2 * mc2 + eV) / (voltage2lambda(volts) * volts)


def scherzer_defocus(lam, cs):
    """Scherzer defocus sqrt(1.5 * Cs * lambda) (optics.py:69-72)."""
    return np.sqrt(1.5) * np.sqrt(cs * lam)


# ---------------------------------------------------------------- geometry
def geometry(xyz, sampling, energy_kev):
    """Simulation geometry as MSA.__init__ derives it (msa.py:96-123).

    xyz: (A,3) float32 positions (x, y, z) in Angstrom.  sampling: (Ny, Nx).
    Returns a dict with dims (float32[3]), kmax, max_ang, pix_size (py, px),
    kpix_size, lam, sigma (all float64).
    """
    xyz = np.asarray(xyz, dtype=np.float32)
    sampling = np.asarray(sampling, dtype=np.int64)
    dims = xyz.max(0) - xyz.min(0)                          # float32, msa.py:110
    kmax = float(np.min(sampling / dims[:2].astype(np.float64)))   # msa.py:114
    lam = float(voltage2lambda(energy_kev * 1e3))
    pix_size = dims[:2][::-1].astype(np.float64) / sampling   # msa.py:121 -> (py, px)
    return dict(dims=dims, kmax=kmax, max_ang=kmax * lam, pix_size=pix_size,
                kpix_size=kmax / sampling, lam=lam,
                sigma=float(sigma_int(energy_kev * 1e3)), sampling=sampling)


def num_slices(dims_z, dz):
    """floor(dims_z / dz) taken in double (msa.py:584)."""
    return int(np.floor(np.float64(dims_z) / dz))


# ---------------------------------------------------------------- potentials
def projected_potential(r, params, eps=BESSEL_EPS):
    """Kirkland projected potential at radius r (msa.py:180-193).

    params: (4,3) rows a, b, c, d.  v = 4 pi^2 a0 e sum a K0(2 pi r sqrt(b) + eps)
    + 2 pi^2 a0 e sum (c/d) exp(-pi^2 r^2 / d).
    """
    r = np.asarray(r, dtype=np.float64)[..., None]
    a, b, c, d = params
    pref = 4 * np.pi ** 2 * A0_BOHR * E_CHARGE
    bessel = pref * a * k0(2 * np.pi * r * np.sqrt(b) + eps)
    gauss = pref / 2.0 * (c / d) * np.exp(-np.pi ** 2 * r ** 2 / d)
    return (bessel + gauss).sum(-1)


def table_shape(pix_size, potential_range=8.0):
    """(Py, Px) = floor(range / pix) per axis (msa.py:142, utils.py:101-108)."""
    return tuple(int(np.floor(potential_range / p)) for p in pix_size)


def atomic_table(params, pix_size, potential_range=8.0, oversample=2):
    """One float32 P_y x P_x table (msa.py:137-146, :180-196; utils.py:101-108).

    Subsample axes are linspace(-range/2, range/2, oversample*P) inclusive; the
    table is the oversample x oversample block mean, shifted so its minimum is 0.
    """
    py, px = table_shape(pix_size, potential_range)
    half = potential_range / 2
    g0 = np.linspace(-half, half, oversample * py)
    g1 = np.linspace(-half, half, oversample * px)
    r = np.sqrt(g0[:, None] ** 2 + g1[None, :] ** 2)
    v = projected_potential(r, params)
    v = v.reshape(py, oversample, px, oversample).mean(axis=(1, 3))
    v = v - v.min()
    return v.astype(np.float32)


def bin_sites(xyz, Z, dz, pix_size, dims):
    """Per-slice site lists under the MSAGPU rule (msa.py:584-604).

    Slice n holds atoms with float32(n*dz) <= z < float32((n+1)*dz).  A site is
    (table index of Z among sorted unique Z, int(y/py), int(x/px)) with
    truncation toward zero.  Returns (list of (k,3) int arrays, unique_Z).
    """
    xyz = np.asarray(xyz, dtype=np.float32)
    Z = np.asarray(Z)
    S = num_slices(dims[-1], dz)
    uniq = np.unique(Z)
    zidx = np.array([int(np.nonzero(uniq == z)[0][0]) for z in Z], dtype=np.int64)
    ypix = (xyz[:, 1].astype(np.float64) / pix_size[0]).astype(np.int32)
    xpix = (xyz[:, 0].astype(np.float64) / pix_size[1]).astype(np.int32)
    out = []
    for n in range(S):
        lo = np.float32(n * dz)
        hi = np.float32((n + 1) * dz)
        m = (xyz[:, 2] >= lo) & (xyz[:, 2] < hi)
        out.append(np.stack([zidx[m], ypix[m], xpix[m]], axis=1).astype(np.int64))
    return out, uniq


def _rintf(v):
    """rintf on a float32 value: round half to even (potential_kernels.cu:30-33)."""
    return int(np.rint(np.float32(v)))


def splat_slices(tables, sites_per_slice, Ny, Nx):
    """Projected potential V[S][Ny][Nx] in FP64 (potential_kernels.cu:21-45).

    For a site with pixel centre (yc, xc): y0 = rintf(yc - Py/2), y1 =
    rintf(yc + Py/2); rows y0 <= row < y1, 0 <= row < Ny, 0 <= row-y0 < Py
    receive tables[z][row-y0][col-x0] (same rule on columns).  Sentinel sites
    (z < 0) are skipped.
    """
    tables = np.asarray(tables, dtype=np.float32)
    _, Py, Px = tables.shape
    V = np.zeros((len(sites_per_slice), Ny, Nx), dtype=np.float64)
    for s, sites in enumerate(sites_per_slice):
        for z, yc, xc in np.asarray(sites).reshape(-1, 3):
            if z < 0:
                continue
            y0 = _rintf(yc - Py * 0.5)
            y1 = _rintf(yc + Py * 0.5)
            x0 = _rintf(xc - Px * 0.5)
            x1 = _rintf(xc + Px * 0.5)
            r_lo, r_hi = max(y0, 0), min(y1, Ny, y0 + Py)
            c_lo, c_hi = max(x0, 0), min(x1, Nx, x0 + Px)
            if r_lo >= r_hi or c_lo >= c_hi:
                continue
            V[s, r_lo:r_hi, c_lo:c_hi] += tables[z, r_lo - y0:r_hi - y0,
                                                 c_lo - x0:c_hi - x0]
    return V


def transmission(V, sigma_f32):
    """t = exp(i sigma V) with NaN -> 1 per component (potential_kernels.cu:48-57)."""
    ph = np.asarray(V, dtype=np.float64) * np.float64(np.float32(sigma_f32))
    re = np.cos(ph)
    im = np.sin(ph)
    re[np.isnan(re)] = 1.0
    im[np.isnan(im)] = 1.0
    return re + 1j * im


# ---------------------------------------------------------------- probe / propagator
def k_axis(n, kmax_f32):
    """double k_j = j*kmax/(n-1) - kmax/2 (probe_kernels.cu:11-16)."""
    km = np.float64(np.float32(kmax_f32))
    j = np.arange(n, dtype=np.float64)
    return j * km / np.float64(n - 1) - km / 2.0


def k_radius(Ny, Nx, kmax_f32):
    ky = k_axis(Ny, kmax_f32)
    kx = k_axis(Nx, kmax_f32)
    return np.sqrt(kx[None, :] * kx[None, :] + ky[:, None] * ky[:, None])


def aperture(Ny, Nx, kmax_f32, k_semi_f32, smooth=True, steepness=80.0):
    """Soft/hard objective aperture (probe_kernels.cu:180-206).

    k is rounded to float32 as the kernels store it; soft: 1/(1+exp(-2*s*(k_semi-k)));
    hard: 1 where float32(k) < float32(k_semi).
    """
    kf = k_radius(Ny, Nx, kmax_f32).astype(np.float32).astype(np.float64)
    ks = np.float64(np.float32(k_semi_f32))
    if smooth:
        return expit(2.0 * steepness * (ks - kf))   # 1/(1+exp(-x)) without overflow far outside
    return (kf < ks).astype(np.float64)


def chi_phase(Ny, Nx, kmax_f32, lam_f32, C1, C3, C5):
    """Aberration phase chi (probe_kernels.cu:208-222), float constants as there."""
    k = k_radius(Ny, Nx, kmax_f32)
    lam = np.float64(np.float32(lam_f32))
    c1 = np.float64(np.float32(-0.5) * np.float32(C1))
    c3 = np.float64(np.float32(0.25) * np.float32(C3))
    c5 = np.float64(np.float32(np.float32(1.0) / np.float32(6.0)) * np.float32(C5))
    kl = k * lam
    return 2.0 * PI_F32 / lam * (c1 * kl ** 2 + c3 * kl ** 4 + c5 * kl ** 6)


def probe_k(Ny, Nx, kmax, k_semi, lam, C1=0.0, C3=0.0, C5=0.0, smooth=True,
            steepness=80.0):
    """psi_k = A(k) * exp(-i chi(k)) on the linspace grid (msa.py:669-715)."""
    A = aperture(Ny, Nx, kmax, k_semi, smooth, steepness)
    chi = chi_phase(Ny, Nx, kmax, lam, C1, C3, C5)
    return A * np.exp(-1j * chi), A


def fftshift_even(a):
    """In-place quadrant swap of fftshift_2d(_stack) for even sizes (probe_kernels.cu:256-304)."""
    return np.fft.fftshift(a, axes=(-2, -1))


def ifft2_u(a):
    """Unnormalised inverse 2D DFT (cuFFT INVERSE)."""
    n = a.shape[-1] * a.shape[-2]
    return np.fft.ifft2(a, axes=(-2, -1)) * n


def fft2_u(a):
    """Unnormalised forward 2D DFT (cuFFT FORWARD)."""
    return np.fft.fft2(a, axes=(-2, -1))


def probe_real(psi_k):
    """Real-space probe and normalization (msa.py:720-728)."""
    psi = fftshift_even(ifft2_u(psi_k))
    norm = float(np.sum(np.abs(psi) ** 2))
    return psi, norm


def probe_stack(psi_k, pos_yx, kmax_f32):
    """Shifted probes (probe_kernels.cu:37-57, :239-253; msa.py:959-964).

    pos_yx: (B,2) float32 (ry, rx) in Angstrom.  ramp = exp(2 pi i kr),
    kr = -kx*rx - ky*ry, evaluated in double.
    """
    Ny, Nx = psi_k.shape
    ky = k_axis(Ny, kmax_f32)
    kx = k_axis(Nx, kmax_f32)
    pos = np.asarray(pos_yx, dtype=np.float32).astype(np.float64).reshape(-1, 2)
    ry = pos[:, 0][:, None, None]
    rx = pos[:, 1][:, None, None]
    kr = -kx[None, None, :] * rx - ky[None, :, None] * ry
    return fftshift_even(ifft2_u(psi_k[None] * np.exp(2j * np.pi * kr)))


def bandwidth_mask(Ny, Nx, kmax_f32, k_cut_f32):
    """1 where float32(k) < float32(k_cut) (hard_aperture, probe_kernels.cu:180-194)."""
    kf = k_radius(Ny, Nx, kmax_f32).astype(np.float32)
    return (kf < np.float32(k_cut_f32)).astype(np.float64)


def propagator_mask(Ny, Nx, kmax_f32, k_cut_f32, lam_f32, dz_f32):
    """PM = exp(-i pi k^2 lambda dz) * mask (probe_kernels.cu:167-178; msa.py:833-839).

    k is the float32-rounded radius the kernel uses; the phase is FP64.
    """
    kf = k_radius(Ny, Nx, kmax_f32).astype(np.float32).astype(np.float64)
    ph = np.pi * kf * kf * np.float64(np.float32(lam_f32)) * np.float64(np.float32(dz_f32))
    return np.exp(-1j * ph) * bandwidth_mask(Ny, Nx, kmax_f32, k_cut_f32)


# ---------------------------------------------------------------- multislice
def multislice(psi0, t, PM, order="ascending"):
    """Per-slice psi <- IFFT_u(PM * FFT(psi * t_i / N^2)) (msa.py:971-981).

    psi0: (B,Ny,Nx) complex; t: (S,Ny,Nx) complex.  order 'descending' follows
    the CPU MSA loop over slices[::-1] (msa.py:346).  Returns exit waves.
    """
    psi = np.array(psi0, dtype=np.complex128, copy=True)
    n = psi.shape[-1] * psi.shape[-2]
    idx = range(t.shape[0]) if order == "ascending" else range(t.shape[0] - 1, -1, -1)
    for i in idx:
        psi = ifft2_u(PM * fft2_u(psi * t[i] / n))
    return psi


def readout(psi_exit, norm):
    """CBED = |sqrt(N^2) FFT(psi/N^2)|^2 / normalization (msa.py:985-990, :905)."""
    n = psi_exit.shape[-1] * psi_exit.shape[-2]
    return np.abs(np.sqrt(n) * fft2_u(psi_exit / n)) ** 2 / norm


def cbed(psi_k, pos_yx, kmax_f32, t, PM):
    """Probe stack -> multislice -> readout, the whole MSAGPU transmit path."""
    psi0 = probe_stack(psi_k, pos_yx, kmax_f32)
    _, norm = probe_real(psi_k)
    return readout(multislice(psi0, t, PM), norm)


# ---------------------------------------------------------------- raster / detectors
def probe_positions(dims, grid_steps=None, probe_step=(0.1, 0.1),
                    probe_range=((0.0, 1.0), (0.0, 1.0)), fraction=0.5,
                    origin=(0, 0)):
    """Probe raster [P][2] = (y, x) (msa.py:253-278).

    grid_steps branch: x over linspace(start[0], stop[0], gs[0]) (slow axis),
    y over linspace(start[1], stop[1], gs[1]) (fast axis), start/stop =
    (0.5 + origin[::-1]*dims -/+ dims*fraction)/2.
    """
    d = np.asarray(dims, dtype=np.float64)[:2]
    org = np.asarray(origin, dtype=np.float64)[::-1]
    if grid_steps is not None:
        gs = np.asarray(grid_steps).astype(np.int64)
        start = (0.5 + org * d - d * fraction) / 2
        stop = (0.5 + org * d + d * fraction) / 2
        ax0 = np.linspace(start[0], stop[0], gs[0])
        ax1 = np.linspace(start[1], stop[1], gs[1])
    else:
        pr = np.asarray(probe_range, dtype=np.float64)
        gs = np.floor(np.diff(pr).flatten() * d / np.asarray(probe_step)).astype(np.int64)
        lo_hi = [(pr[i] - 0.5) * d[i] for i in range(2)]
        ax0 = np.linspace(lo_hi[0][0], lo_hi[0][1], gs[0])
        ax1 = np.linspace(lo_hi[1][0], lo_hi[1][1], gs[1])
    X, Y = np.meshgrid(ax0, ax1, indexing="ij")
    return np.stack([Y.ravel(), X.ravel()], axis=1), gs


def pixel_disc(shape, radius):
    """heaviside(max(shape)*radius - r, 0) on the integer grid (msa.py:245-251)."""
    g0, g1 = np.meshgrid(np.arange(-(shape[0] // 2), shape[0] // 2),
                         np.arange(-(shape[1] // 2), shape[1] // 2), indexing="ij")
    r = np.sqrt(g0 ** 2 + g1 ** 2)
    return (max(shape) * radius - r > 0).astype(np.float64)


def detector_mask_reference(sampling, kmax, lam, inner, outer):
    """Annulus of integrate_cbed's default branch (msa.py:426-435).

    Radius fraction = angle / (kmax*lam) / 2 (kpix*lambda*sampling = kmax*lambda).
    """
    max_ang = kmax * lam
    assert inner < outer < max_ang
    shape = tuple(int(s) for s in sampling)
    ri = inner / max_ang / 2
    ro = outer / max_ang / 2
    return np.logical_xor(pixel_disc(shape, ro) > 0, pixel_disc(shape, ri) > 0).astype(np.float64)


def integrate_cbed(cbeds, grid_steps, det):
    """sum(CBED * detector) per raster point, shaped grid_steps (msa.py:436-439)."""
    shape = list(np.asarray(grid_steps).astype(int)) + list(cbeds.shape[-2:])
    return np.sum(cbeds.reshape(shape) * det, axis=(-2, -1))



# ---------------------------------------------------- alternate slice sources
def integrate_slices(pot, grid, slice_thickness):
    """Slices of a sampled 3D potential (MSA.integrate_potential_slices,
    msa.py:373-407): crop the plane to the square min(Nx, Ny) (msa.py:386-390),
    S = int(z extent // dz), d = Z // S, crop z to S*d samples when d differs
    from ceil(Z/S) (msa.py:393-399), split into S chunks of d samples, trapz
    over each chunk with its own coordinates (no interval spans two chunks),
    negate (msa.py:402-405).

    pot [Z][Ny][Nx] (spatial order z, y, x), grid = (z, y, x) coordinate arrays.
    Returns (V [S][M][M] float64, dict(num_slices, d, z_used, max_dim)).
    """
    z, y, x = (np.asarray(c, dtype=np.float64) for c in grid)
    S = int((z[-1] - z[0]) // slice_thickness)
    M = min(x.size, y.size)
    pot = np.asarray(pot, dtype=np.float64)[:, :M, :M]
    d = int(pot.shape[0] // S)
    if d != int(np.ceil(pot.shape[0] / S)):
        pot = pot[: d * S]
        z = z[: d * S]
    V = np.empty((S, M, M))
    for s in range(S):
        c, zc = pot[s * d:(s + 1) * d], z[s * d:(s + 1) * d]
        acc = np.zeros((M, M))
        for j in range(1, d):   # numpy's trapz terms, summed in z order
            acc = acc + (zc[j] - zc[j - 1]) * (c[j] + c[j - 1]) / 2.0
        V[s] = -acc
    return V, {"num_slices": S, "d": d, "z_used": d * S, "max_dim": M}


def integration_params(slices_shape, grid, lam, max_angle=None):
    """Simulation parameters after integrate_potential_slices (msa.py:409-421):
    dims from the coordinate maxima of the full grids, sampling from the slice
    shape (or from kmax * dims when max_angle is set), kmax, max_ang, pix_size
    and kpix_size."""
    z, y, x = (np.asarray(c, dtype=np.float64) for c in grid)
    dims = np.array([x.max(), y.max(), z.max()])
    if max_angle is None:
        sampling = np.array(slices_shape[1:])
        kmax = np.min(sampling / dims[:2])
        max_ang = kmax * lam
    else:
        max_ang = max_angle
        kmax = max_ang / lam
        sampling = np.floor(kmax * dims[:2]).astype(np.int64)[::-1]
    pix_size = dims[:2][::-1] / sampling
    return {"dims": dims, "sampling": sampling, "kmax": kmax, "max_ang": max_ang, "pix_size": pix_size,
            "kpix_size": kmax / sampling}


def thin_object(V, probes):
    """|FFT2(probes * exp(i sum_s V_s))| (thin_object_approx, utils.py:130-135)."""
    psi = np.asarray(probes, dtype=np.complex128) * np.exp(1j * np.asarray(V, dtype=np.float64).sum(0))
    return np.abs(np.fft.fft2(psi))
