"""namsa_amd: MI355X-native multislice electron scattering (drop-in for nlaanait/namsa MSAGPU).

Public surface mirrors the reference package (namsa/__init__.py:2-5):
``MSAGPU`` (namsa/msa.py:529-992) plus the optics helpers.  Compute runs in
hand-written gfx950 HIP kernels (libnamsa_hip.so, C ABI in include/namsa_hip.h)
called through ctypes; build them with ``python -m namsa_amd.build``.
"""
from ._lib import NamsaError
from .msa import DEFAULT_PROBE, XYZ_dtype, MSAGPU, kirkland_params
from .optics import scherzer_params, sigma_int, voltage2Lambda

__all__ = ["MSAGPU", "NamsaError", "DEFAULT_PROBE", "XYZ_dtype", "kirkland_params",
           "voltage2Lambda", "sigma_int", "scherzer_params"]
