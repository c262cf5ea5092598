"""Training records for 4D-STEM datasets: the step after the slice loop in
the reference's dataset scripts (scripts/summit_scripts/utils.py:40-118,
scripts/summit_scripts/sim_prod.py:119-137, :186-197).

process_cbed and process_potential run on the GPU (namsa_cbed_records,
namsa_potential_record). They accept host arrays or device tensors, so a
simulation's CBEDs can be encoded where they lie; only the half-precision
records then cross PCIe. write_lmdb, write_lmdb_headers and write_h5 keep the
reference's key names, byte layouts and group layout. They write through any
object with the lmdb transaction's put(key, value), or h5py's
create_group / create_dataset / attrs, since neither package is a dependency
here.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import NamsaError, call, ptr, stream_handle

RECORD_NAMES = ("2d_potential_", "cbed_")


def _torch():
    import torch

    return torch


def _device_f32(x, device):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            x = x.to(device)
        return x.contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)


def process_cbed(cbed, normalize=True, scale=(-1, 1), fp16=False, new_shape=None, device=None, as_tensor=False,
                 moments=None):
    """Standardise each CBED of [P][Ny][Nx] by its own mean and population
    deviation and return float16 (utils.py:104-118). The reference's keyword
    arguments are accepted and, as there, unused.

    cbed: numpy array or torch tensor (a CUDA tensor is read where it lies,
    without a host copy; the input is not modified). moments: the [P][2] (sum, sum of squares) of the same
    CBEDs, as MSAGPU.multislice(moments=True) keeps them; then the encoder
    reads the CBEDs once instead of twice. Returns numpy float16, or the
    device tensor with as_tensor=True.
    """
    torch = _torch()
    _lib.load()
    dev = (cbed.device if isinstance(cbed, torch.Tensor) and cbed.is_cuda
           else torch.device("cuda", 0) if device is None else torch.device(device))
    x = _device_f32(cbed, dev)
    if x.dtype != torch.float32:
        x = x.float()
    if x.dim() != 3:
        raise NamsaError(f"process_cbed expects [P][Ny][Nx], got shape {tuple(x.shape)}")
    P, Ny, Nx = (int(v) for v in x.shape)
    out = torch.empty((P, Ny, Nx), dtype=torch.float16, device=dev)
    if moments is not None:
        mom = (moments if isinstance(moments, torch.Tensor) else torch.from_numpy(np.asarray(moments)))
        mom = mom.to(device=dev, dtype=torch.float64).contiguous()
        if tuple(mom.shape) != (P, 2):
            raise NamsaError(f"moments must be [{P}][2], got {tuple(mom.shape)}")
        with torch.cuda.device(dev):
            call("namsa_cbed_records_from_moments", ptr(x), ptr(mom), P, Ny, Nx, ptr(out), stream_handle())
        return out if as_tensor else out.cpu().numpy()
    ws = torch.empty(max(1, int(_lib.load().namsa_records_workspace_doubles(P, Ny, Nx))), dtype=torch.float64,
                     device=dev)
    with torch.cuda.device(dev):
        call("namsa_cbed_records", ptr(x), P, Ny, Nx, ptr(out), ptr(ws), stream_handle())
    return out if as_tensor else out.cpu().numpy()


def potential_record_shape(Ny, Nx):
    rows, cols = ctypes.c_int(), ctypes.c_int()
    call("namsa_potential_record_shape", int(Ny), int(Nx), ctypes.byref(rows), ctypes.byref(cols))
    return rows.value, cols.value


def process_potential(pot_slices, normalize=True, mask=None, sampling=None, scale=(-1, 1), expand_dim=True,
                      fp16=False, sigma=1.2, truncate=4.0, device=None, as_tensor=False):
    """The projected-potential record of a set of transmission slices
    (utils.py:74-102): mean over slices of imag(t), Gaussian filter (sigma 1.2),
    the central crop [Ny/4, 3Nx/4) on both axes, standardised (deviation
    floored at 1/sqrt(size)), float16 of shape [1][rows][cols]. As in the
    reference, mask, sampling, scale and the flags are accepted and unused.
    """
    torch = _torch()
    _lib.load()
    if isinstance(pot_slices, torch.Tensor) and pot_slices.is_cuda:
        dev, t = pot_slices.device, pot_slices.contiguous()
    else:
        dev = torch.device("cuda", 0) if device is None else torch.device(device)
        t = torch.from_numpy(np.ascontiguousarray(pot_slices, dtype=np.complex64)).to(dev)
    if t.dtype != torch.complex64:
        t = t.to(torch.complex64)
    if t.dim() == 2:
        t = t[None]
    S, Ny, Nx = (int(v) for v in t.shape)
    rows, cols = potential_record_shape(Ny, Nx)
    out = torch.empty((1, rows, cols), dtype=torch.float16, device=dev)
    ws = torch.empty(int(_lib.load().namsa_potential_record_workspace_bytes(Ny, Nx)), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        call("namsa_potential_record", ptr(t), S, Ny, Nx, float(sigma), float(truncate), ptr(out), ptr(ws),
             stream_handle())
    return out if as_tensor else out.cpu().numpy()


# --- containers ---------------------------------------------------------------------------------------------------


def lmdb_header_entries(input_shape=(64, 256, 256), output_shape=(1, 128, 128), record_names=RECORD_NAMES):
    """The header records of sim_prod.py:190-197, in order. The shapes are
    int64 bytes. 'output_name' holds record_names[0] and 'input_name'
    record_names[1]. The last entry is b'header_entries' -> bytes(n), which
    is n zero bytes for n header entries."""
    entries = [(b"input_dtype", b"float16"),
               (b"input_shape", np.asarray(input_shape, dtype=np.int64).tobytes()),
               (b"output_shape", np.asarray(output_shape, dtype=np.int64).tobytes()),
               (b"output_dtype", b"float16"),
               (b"output_name", bytes(record_names[0], "ascii")),
               (b"input_name", bytes(record_names[1], "ascii"))]
    entries.append((b"header_entries", bytes(len(entries))))
    return entries


def write_lmdb_headers(txn, input_shape=(64, 256, 256), output_shape=(1, 128, 128), record_names=RECORD_NAMES):
    for key, val in lmdb_header_entries(input_shape, output_shape, record_names):
        txn.put(key, val)


def write_lmdb(txn, idx, cbed, potential, record_names=RECORD_NAMES, params=None):
    """utils.py:60-72: b'2d_potential_<idx>' -> the potential record's raw
    bytes, b'cbed_<idx>' -> the CBED record's raw bytes (C order)."""
    pot = np.ascontiguousarray(_host(potential))
    cb = np.ascontiguousarray(_host(cbed))
    txn.put(bytes("%s%s" % (record_names[0], format(idx)), "ascii"), pot.tobytes())
    txn.put(bytes("%s%s" % (record_names[1], format(idx)), "ascii"), cb.tobytes())


def write_h5(h5group, cbed, potential, params):
    """utils.py:40-58: a new group 'sample_<n>' (n = items already in the
    group) holding datasets 'CBED' and 'potential'. The parameters become
    group attributes; probe_params are flattened, with the aberration
    dictionary stored as 'aberrations' = [C1, C3, C5]."""
    g = h5group.create_group("sample_%d" % len(h5group.items()))
    g.create_dataset("CBED", data=_host(cbed))
    g.create_dataset("potential", data=_host(potential))
    for key, val in params.items():
        if key == "probe_params":
            for pk, pv in val.items():
                if pk == "aberration_dict":
                    g.attrs["aberrations"] = np.array([pv["C1"], pv["C3"], pv["C5"]])
                else:
                    g.attrs[pk] = pv
        else:
            g.attrs[key] = val
    return g


def write_4d_cbed(h5file, cbed, total_num_probes=None, offset=0, name="4D_CBED"):
    """The MSAMPI layout (namsa/msa.py:1044-1051): one dataset '4D_CBED' of
    [total probes][Ny][Nx] in the CBED dtype; this process's rows go at
    `offset`. The dataset is created when absent."""
    cb = _host(cbed)
    total = cb.shape[0] if total_num_probes is None else int(total_num_probes)
    if name in h5file:
        dset = h5file[name]
    else:
        dset = h5file.create_dataset(name, (total, cb.shape[1], cb.shape[2]), dtype=cb.dtype)
    dset[offset:offset + cb.shape[0], :, :] = cb
    return dset


def _host(x):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
