"""Electron-optics closed forms used by the host side of the simulation.

Same formulas as the reference's namsa/optics.py (they are physics constants,
not implementation choices): relativistic wavelength (optics.py:4-12),
interaction parameter (optics.py:100-108), Scherzer defocus (optics.py:69-72).
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


def scherzer_params(Lambda, Cs):
    """(defocus, aperture) at Scherzer conditions.

    The aperture keeps the reference's operator precedence, (6*lambda/Cs)**1/4
    == (6*lambda/Cs)/4; callers only use the defocus.
    """
    defocus = np.sqrt(1.5) * np.sqrt(Cs * Lambda)
    aperture = (6 * Lambda / Cs) ** 1 / 4
    return defocus, aperture
