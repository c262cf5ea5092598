"""ctypes binding of libnamsa_hip.so (the C ABI declared in include/namsa_hip.h).

This replaces the reference's kernel FFI (namsa/cuda_kernels.py:8-56: Jinja2
render -> pycuda SourceModule JIT -> get_function) and its scikit-cuda cuFFT
plans.  The library is built ahead of time by ``python -m namsa_amd.build``;
if it is missing, every call raises ``NamsaError`` -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libnamsa_hip.so")
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "namsa_hip.h"
ABI_VERSION = 3

OUT_CBED = 0
OUT_INCIDENT = 1
OUT_EXIT = 2
PM_TILED = 0x100


class NamsaError(RuntimeError):
    """A libnamsa_hip call failed, or the library is not available."""


_vp, _i, _f, _d, _ll, _sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_longlong, C.c_size_t

# name -> (restype, argtypes); must match include/namsa_hip.h
SIGNATURES = {
    "namsa_abi_version": (_i, []),
    "namsa_last_error": (C.c_char_p, []),
    "namsa_init": (_i, [_i]),
    "namsa_fft_supported": (_i, [_i]),
    "namsa_atomic_tables": (_i, [_vp, _i, _i, _i, _d, _i, _vp, _vp]),
    "namsa_build_slices": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _f, _vp, _vp]),
    "namsa_probe_k": (_i, [_i, _i, _f, _f, _f, _f, _f, _f, _i, _f, _vp, _vp, _vp]),
    "namsa_propagator": (_i, [_i, _i, _f, _f, _f, _f, _vp, _vp, _vp]),
    "namsa_fft2d": (_i, [_vp, _i, _i, _i, _i, _vp]),
    "namsa_norm2": (_i, [_vp, _ll, _vp, _vp]),
    "namsa_det_workspace_doubles": (_sz, [_i, _i, _i, _i]),
    "namsa_integrate": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _vp]),
    "namsa_multislice": (_i, [_vp, _vp, _i, _i, _i, _f, _vp, _i, _vp, _i, _i, _i, _i, _f, _vp, _vp,
                              _vp, _i, _vp, _vp, _vp]),
    "namsa_multislice_moments": (_i, [_vp, _vp, _i, _i, _i, _f, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp, _vp, _i,
                                      _vp, _vp, _vp, _vp, _vp]),
    "namsa_multislice_noalloc": (_i, [_vp, _vp, _i, _i, _i, _f, _vp, _i, _vp, _i, _i, _i, _f, _vp, _vp, _i,
                                      _vp, _vp]),
    "namsa_sched_count": (_i, [_i, _i]),
    "namsa_sched_name": (C.c_char_p, [_i, _i, _i]),
    "namsa_sched_get": (_i, [_i, _i]),
    "namsa_sched_set": (_i, [_i, _i, _i]),
    "namsa_probe_init": (_i, [_vp, _vp, _i, _i, _i, _f, _vp, _vp]),
    "namsa_slice_rows": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "namsa_slice_cols": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "namsa_wave_tiled": (_i, [_i, _i]),
    "namsa_slice_rows_tiled": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "namsa_slice_cols_tiled": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "namsa_tile_plane": (_i, [_vp, _vp, _i, _i, _vp]),
    "namsa_build_potential": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _vp, _vp]),
    "namsa_gemm_bf16": (_i, [_vp, _ll, _ll, _ll, _vp, _ll, _ll, _ll, _vp, _ll, _ll, _ll, _i, _i, _i, _i, _i,
                             _f, _f, _vp, _i, _i, _vp, _vp]),
    "namsa_fill": (_i, [_vp, _ll, _f, _vp]),
    "namsa_gelu_backward": (_i, [_vp, _vp, _ll, _vp]),
    "namsa_adam": (_i, [_vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _i, _f, _vp]),
    "namsa_row_fft": (_i, [_vp, _i, _i, _i, _i, _f, _vp]),
    "namsa_precise_slices": (_i, [_i]),
    "namsa_sched_precise": (_i, [_i, _i, _i]),
    "namsa_integrate_slices": (_i, [_vp, _vp, _i, _i, _i, _i, _d, _vp, _vp, _vp]),
    "namsa_thin_object": (_i, [_vp, _i, _vp, _i, _i, _i, _vp, _vp, _vp]),
    "namsa_records_workspace_doubles": (_ll, [_i, _i, _i]),
    "namsa_cbed_records": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "namsa_cbed_records_from_moments": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "namsa_potential_record_shape": (_i, [_i, _i, _vp, _vp]),
    "namsa_potential_record_workspace_bytes": (_ll, [_i, _i]),
    "namsa_potential_record": (_i, [_vp, _i, _i, _i, _d, _d, _vp, _vp, _vp]),
    "namsa_spectral_pack": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _vp, _f, _f, _vp, _f, _f, _i,
                                 _vp, _vp]),
    "namsa_spectral_loss": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _f, _vp, _vp, _vp]),
    "namsa_spectral_apply": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _f, _vp, _vp]),
    "namsa_real_to_complex": (_i, [_vp, _i, _i, _i, _f, _i, _vp, _vp]),
    "namsa_fno_lift": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp]),
    "namsa_fno_layer": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp]),
    "namsa_modes_combine": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "namsa_modes_expand": (_i, [_vp, _vp, _i, _i, _i, _vp]),
}

_lib = None


def load() -> C.CDLL:
    """Load the shared library once and bind every declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NamsaError(f"{LIB_PATH} is not built; run `python -m namsa_amd.build`")
    # torch ships its own HIP runtime; it must be the one in the process before
    # the library binds libamdhip64 (which then resolves to that same copy, so
    # torch's streams and allocations are valid here).  Loading the library
    # first would bind the system runtime and leave torch without devices.
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    ver = lib.namsa_abi_version()
    if ver != ABI_VERSION:
        raise NamsaError(f"libnamsa_hip ABI {ver} != expected {ABI_VERSION}")
    _lib = lib
    return lib


def call(name: str, *args) -> int:
    """Invoke an int-returning entry point; raise NamsaError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.namsa_last_error().decode(errors="replace")
        raise NamsaError(f"{name} failed ({rc}): {msg}")
    return rc


def declared_functions(header: Path = HEADER_PATH) -> list[str]:
    """Names of the functions the C header declares."""
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(namsa_\w+)\s*\(", text)))


def ptr(t) -> C.c_void_p:
    """Device pointer of a torch tensor (or None)."""
    return C.c_void_p(0 if t is None else t.data_ptr())


def stream_handle(stream=None) -> C.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
