"""Electron-optics closed forms used by the host side of the simulation.

Same formulas as the reference's namsa/optics.py (they are physics constants,
not implementation choices): relativistic wavelength (optics.py:4-12),
interaction parameter (optics.py:100-108), Scherzer defocus (optics.py:69-72),
and the parameter-choice helpers the reference's dataset scripts call
(optics.py:15-66, :74-97, :110-114).
"""
from __future__ import annotations

import numpy as np
from scipy.constants import elementary_charge, electron_mass, speed_of_light

_MC2 = electron_mass * speed_of_light ** 2


def voltage2Lambda(voltage):
    """Electron wavelength in Angstrom for an accelerating voltage in volts."""
    return 12.2643 / np.sqrt(voltage + 0.97845e-6 * voltage * voltage)


def sigma_int(voltage):
    """Interaction parameter in rad / (V Angstrom)."""
    e_v = elementary_charge * voltage
    return 2 * np.pi * (_MC2 + e_v) / (2 * _MC2 + e_v) / (voltage2Lambda(voltage) * voltage)


def scherzer_defocus(Lambda, Cs):
    """Scherzer defocus sqrt(1.5 Cs lambda) alone: the defocus term of
    scherzer_params, without its aperture (which divides by Cs)."""
    return np.sqrt(1.5) * np.sqrt(Cs * Lambda)


def scherzer_params(Lambda, Cs):
    """(defocus, aperture) at Scherzer conditions.

    The aperture keeps the reference's operator precedence, (6*lambda/Cs)**1/4
    == (6*lambda/Cs)/4; callers only use the defocus.
    """
    defocus = scherzer_defocus(Lambda, Cs)
    aperture = (6 * Lambda / Cs) ** 1 / 4
    return defocus, aperture


def transverse_width(Lambda, theta_s):
    """Transverse coherence width lambda / (2 pi theta_s) in Angstrom, theta_s
    the source's convergence semi-angle in rad (Zernike-Van Cittert)."""
    return Lambda / (2 * np.pi * theta_s)


def no_overlap_dhkl(Lambda, C_3=1):
    """Largest d-spacing (Angstrom) whose CBED orders do not overlap for a
    diffraction-limited probe: 0.7874 Cs^(1/4) lambda^(3/4), Cs given in mm."""
    return 0.7874 * (C_3 * 1.0e7) ** 0.25 * Lambda ** 0.75


def semi_angle_limit(Lambda, C_3=1):
    """Convergence semi-angle (mrad) of a diffraction-limited probe:
    1.51 Cs^(-1/4) lambda^(1/4), Cs given in mm."""
    return 1.51 * (C_3 * 1.0e7) ** -0.25 * Lambda ** 0.25 * 1e3


def overlap_params(overlap, d_hkl, Lambda):
    """(aperture in mrad, Cs, defocus in Angstrom) for a convergence angle of
    `overlap` Bragg angles at Scherzer focus (three-beam overlap). Cs is
    returned as the Angstrom value times 1e-6, as the reference does; that is
    ten times its value in mm."""
    theta_c = overlap * Lambda / (2 * d_hkl)
    C_3 = (theta_c / 1.51) ** -4 * Lambda
    defocus = -1.15 * C_3 ** -0.25 * Lambda ** -0.75
    return theta_c * 1e3, C_3 * 1.0e-6, defocus


def spherical_phase_error(k_rad, Lambda, scherzer, **kwargs):
    """exp(-i chi(k)) with chi = 2 pi / lambda (-C1 (k lambda)^2 / 2 +
    C3 (k lambda)^4 / 4 + C5 (k lambda)^6 / 6); scherzer=True takes C1 from
    Scherzer's condition when C3 > 1 (optics.py:74-97). Host numpy; the GPU
    probe former evaluates the same chi in FP64."""
    C1, C3, C5 = kwargs["C1"], kwargs["C3"], kwargs["C5"]
    if scherzer:
        if C3 > 1.0:
            C1 = scherzer_params(Lambda, C3)[0]
        else:
            print("Spherical Aberration is too small!! Not using Scherzer condition.")
    q = np.asarray(k_rad) * Lambda
    chi = 2 * np.pi / Lambda * (-0.5 * C1 * q ** 2 + 0.25 * C3 * q ** 4 + C5 * q ** 6 / 6)
    return np.exp(-1j * chi)


def cal_probe_size(Lambda, semi_angle, scherzer=False):
    """Diffraction-limited probe size 0.61 lambda / semi_angle (Angstrom); like
    the reference, no value (None) for scherzer=True."""
    if scherzer:
        return None
    return 0.61 * Lambda / semi_angle
