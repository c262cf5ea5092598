"""MSAGPU-compatible multislice simulation on MI355X.

Drop-in for the reference class ``MSAGPU`` (namsa/msa.py:529-992) and the
parts of ``MSA`` / ``MSAHybrid`` it inherits (geometry msa.py:66-127, atomic
potentials :137-146, probe raster :253-278, memory planner :445-461, detector
integration :426-439).  Same constructor, method names, argument meaning,
call order and result attributes; the work runs in hand-written gfx950 HIP
kernels (libnamsa_hip.so) reached through ctypes, with torch tensors as the
device buffers and the current torch stream as the HIP stream.

Differences from the reference, each deliberate:
  * sampling is a runtime argument; supported sizes are those with a compiled
    transform schedule (namsa_fft_supported).
  * the probe ramp phase is evaluated in FP64 (the reference forms it in FP32,
    about 1e-4 rad of phase error at 30 Angstrom offsets);
  * padded sites are never read (the reference reads atom_pot_stack[-1] for
    them, namsa/potential_kernels.cu:25-41);
  * batch_size only splits the work, it never reorders probe positions
    (the reference scrambles them, msa.py:773);
  * debye_waller jitter takes an optional numpy Generator for reproducibility.
"""
from __future__ import annotations

import os
from pathlib import Path
from warnings import warn

import numpy as np

from . import _lib
from ._lib import NamsaError, call, ptr, stream_handle
from .optics import scherzer_defocus, sigma_int, voltage2Lambda

XYZ_dtype = [("atomic_number", "i"), ("x", "f"), ("y", "f"), ("z", "f"), ("occ", "f"), ("DW", "f")]
KIRKLAND_PATH = Path(__file__).resolve().parent / "data" / "kirkland_params.npy"
kirkland_params = np.load(KIRKLAND_PATH)  # float64 (103, 4, 3), rows a, b, c, d

DEFAULT_PROBE = {"smooth_apert": True, "apert_smooth": 50, "spherical_phase": True,
                 "aberration_dict": {"C1": 0.0, "C3": 0.0, "C5": 0.0}, "scherzer": True}
GPU_APERTURE_STEEPNESS = 80.0  # fixed in soft_aperture, namsa/probe_kernels.cu:203-204


LLC_BYTES = 256 << 20  # MI355X last-level (Infinity) cache


def _torch():
    import torch

    return torch


# Share of the last-level cache that one probe group's waves should fill, per
# square grid: the fastest of 0.5-1.5x measured on MI355X for the slice loop
# (bandwidth 1/3).  Larger groups amortise the fixed cost of each pass
# launch; past the cache the wave traffic spills to HBM.  At 512^2 the optimum
# (160 probes) lies past the cache and 176 is 6% slower; at 2048^2 the slice
# and propagator planes (64 MB) also compete for the cache.
_GROUP_SHARE = {128: 0.75, 192: 1.0, 256: 0.875, 384: 1.0, 512: 1.25, 768: 1.0, 1024: 1.0, 2048: 0.75}


def cache_group(Ny, Nx):
    """Probes per group: complex64 waves filling the measured best share of
    the last-level cache (3/4 for grids without a measurement).

    Pushing such a group through all slices before starting the next keeps the
    slice loop's wave traffic in the cache instead of HBM.  The autotuner's
    cached choice (namsa_amd.autotune.tune_group) overrides this default.
    """
    from . import autotune

    g = autotune.cached_group(Ny, Nx)
    if g:
        return g
    share = _GROUP_SHARE.get(Nx, 0.75) if Ny == Nx else 0.75
    return max(1, int(LLC_BYTES * share // (8 * Ny * Nx)))


def phase_plan(P, max_probes, group, sink=False, batch_size=None):
    """How multislice splits P probe positions: (chunk, resident).

    chunk: probes per call of the fused driver, at most ``max_probes`` (the
    memory planner's bound, msa.py:445-461) and ``batch_size``; with a sink
    and no batch_size, a multiple of the probe group giving about eight
    chunks, so copies overlap compute.  resident: the whole output stays on
    the device (no sink and every position fits at once); otherwise the run
    goes in phases of ``chunk`` positions, in order (the reference's phases,
    msa.py:843-907), with device memory for two chunks of output only.
    """
    P, group = int(P), max(1, int(group))
    chunk = max(1, min(int(max_probes), P))
    if sink and not batch_size:
        chunk = min(chunk, group * max(1, -(-P // group) // 8))
    if batch_size:
        chunk = min(chunk, int(batch_size))
    return chunk, (not sink) and P <= int(max_probes)


class _ChunkStreamer:
    """Moves finished chunks of a phased multislice off the device.

    The driver writes chunk k into device slot k % n_slots; this object copies
    it on a second stream (default priority) to pinned host memory -- the
    caller's full host output, or double-buffered staging handed to a sink one
    chunk behind the compute stream -- so the copy and the sink overlap the
    next chunk's slices.  With moments, each chunk is first encoded into
    float16 records on the compute stream (namsa_cbed_records_from_moments).
    The compute stream waits for a slot's copy before writing into it again.
    """

    def __init__(self, Ny, Nx, dtype, device, chunk, n_slots, sink=None, host_out=None, moments=None):
        torch = _torch()
        self.Ny, self.Nx, self.sink, self.host_out, self.moments = Ny, Nx, sink, host_out, moments
        stage = torch.float16 if moments is not None else dtype
        self.host = ([torch.empty((chunk, Ny, Nx), dtype=stage, pin_memory=True) for _ in range(2)]
                     if sink is not None else None)
        self.enc = ([torch.empty((chunk, Ny, Nx), dtype=torch.float16, device=device) for _ in range(n_slots)]
                    if moments is not None else None)
        self.copy_stream = torch.cuda.Stream(device=device)
        self.copied = [None] * n_slots      # event: the slot's device data has left the device
        self.done = [None, None]            # event: the staging slot's copy finished
        self.pending = None
        self.k = 0

    def claim(self, slot):
        """Make the compute stream wait until device slot `slot` may be overwritten."""
        if self.copied[slot] is not None:
            _torch().cuda.current_stream().wait_event(self.copied[slot])

    def push(self, src, start, B, slot):
        torch = _torch()
        if self.enc is not None:
            call("namsa_cbed_records_from_moments", ptr(src), ptr(self.moments[start]), B, self.Ny, self.Nx,
                 ptr(self.enc[slot]), stream_handle())
            src = self.enc[slot][:B]
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(src.device))
        self.copy_stream.wait_event(ready)
        if self.sink is None:
            with torch.cuda.stream(self.copy_stream):
                self.host_out[start:start + B].copy_(src, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
            self.copied[slot] = ev
            return
        hs = self.k % 2
        self.k += 1
        if self.done[hs] is not None:      # the staging slot's previous chunk was delivered already
            self.done[hs].synchronize()
        with torch.cuda.stream(self.copy_stream):
            self.host[hs][:B].copy_(src, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(self.copy_stream)
        self.copied[slot] = ev
        self.done[hs] = ev
        if self.pending is not None:
            self._deliver(*self.pending)
        self.pending = (hs, start, B)

    def _deliver(self, hs, start, B):
        self.done[hs].synchronize()
        self.sink(start, self.host[hs][:B].numpy())

    def finish(self):
        if self.pending is not None:
            self._deliver(*self.pending)
            self.pending = None
        self.copy_stream.synchronize()


class MSAGPU:
    """Multislice on one MI355X behind the reference MSAGPU interface."""

    def __init__(self, energy, semi_angle, supercell, sampling=np.array([512, 512]),
                 max_angle=None, verbose=False, debug=False, output_dir="", debye_waller=True,
                 rng=None):
        self.E = energy
        self.Lambda = voltage2Lambda(self.E * 1e3)
        self.semi_ang = semi_angle
        self.verbose = verbose
        self.debug = debug
        self.output_dir = output_dir
        self.debye_waller = debye_waller
        arr = self._load_supercell(supercell)
        self.supercell_xyz = np.column_stack([arr["x"], arr["y"], arr["z"]]).astype(np.float32)
        self.supercell_Z = np.asarray(arr["atomic_number"]).astype(np.int64)
        self.supercell_dw = np.asarray(arr["DW"], dtype=np.float32)
        self.supercell_occ = np.asarray(arr["occ"], dtype=np.float32)
        if self.debye_waller:  # per-element Gaussian displacement of x, y and z (msa.py:101-107)
            gen = rng if rng is not None else np.random
            for Z in np.unique(self.supercell_Z):
                idx = np.nonzero(self.supercell_Z == Z)[0]
                dw = float(self.supercell_dw[idx][0])
                self.supercell_xyz[idx] += (dw * gen.standard_normal(size=(len(idx), 3))).astype(np.float32)

        self.dims = self.supercell_xyz.max(0) - self.supercell_xyz.min(0)
        self.max_angle = max_angle
        dims64 = self.dims[:2].astype(np.float64)
        if max_angle is None:
            self.sampling = np.asarray(sampling, dtype=np.int64)
            self.kmax = float(np.min(self.sampling / dims64))
            self.max_ang = self.kmax * self.Lambda
        else:
            self.max_ang = max_angle
            self.kmax = self.max_ang / self.Lambda
            self.sampling = np.floor(self.kmax * dims64).astype(np.int64)[::-1]
        self.pix_size = self.dims[:2][::-1].astype(np.float64) / self.sampling
        self.kpix_size = self.kmax / self.sampling
        self.sigma = float(sigma_int(self.E * 1e3))
        self.device = None
        self.vars = []
        self.probe_dict = None
        self.num_slices = 0
        self._t = None
        self._probes_dev = None
        self._probes_host = None
        self._slices_host = None
        self.print_verbose(
            "Simulation Parameters:\nSupercell dimensions xyz:%s (A)\nReal, Reciprocal space pixel "
            "sizes:%s A, %s 1/A\nMax angle: %2.2f (rad)\nSampling: %s pixels,\nThermal Effects: %s"
            % (np.round(self.dims, 3), np.round(self.pix_size, 3), np.round(self.kpix_size, 3),
               self.max_ang, self.sampling, self.debye_waller))

    # ------------------------------------------------------------------ helpers
    @staticmethod
    def _load_supercell(supercell):
        if isinstance(supercell, np.ndarray):
            for field in ("x", "y", "z", "occ", "DW", "atomic_number"):
                if field not in (supercell.dtype.names or ()):
                    raise ValueError("Supercell numpy array does not have required fields.")
            return np.copy(supercell)
        if isinstance(supercell, (str, os.PathLike)):
            if not os.path.exists(supercell):
                raise FileNotFoundError("Path %s was not found." % supercell)
            return np.genfromtxt(supercell, skip_footer=1, skip_header=2, dtype=XYZ_dtype)
        raise TypeError("supercell format was not recognized.")

    def print_debug(self, *args, **kwargs):
        if self.debug:
            print(*args, **kwargs)

    def print_verbose(self, *args, **kwargs):
        if self.verbose:
            print(*args, **kwargs)

    def _dev(self):
        if self.device is None:
            self.setup_device(0)
        return self.device

    @property
    def _grid(self):
        return int(self.sampling[0]), int(self.sampling[1])

    def _check_grid(self):
        lib = _lib.load()
        Ny, Nx = self._grid
        if not (lib.namsa_fft_supported(Ny) and lib.namsa_fft_supported(Nx)):
            raise NamsaError(f"sampling {Ny}x{Nx} has no compiled transform schedule")
        if self._t is not None and tuple(self._t.shape[1:]) != (Ny, Nx):
            raise NamsaError(f"potential slices are {tuple(self._t.shape[1:])}, the sampling is {Ny}x{Nx}")

    # ------------------------------------------------------------------ device
    def setup_device(self, gpu_id=0):
        """Select the GPU and bring up the HIP runtime (msa.py:535-553)."""
        torch = _torch()
        if not torch.cuda.is_available():
            raise NamsaError("no HIP device is visible; namsa_amd runs only on the GPU")
        torch.cuda.set_device(gpu_id)
        call("namsa_init", int(gpu_id))
        self.device = torch.device("cuda", gpu_id)
        self.ctx = self.device
        return self.device

    @staticmethod
    def clean_up(ctx=None, vars=None):
        """Drop references to device buffers (msa.py:555-566)."""
        if vars is not None:
            vars.clear()

    # ------------------------------------------------------------------ potentials
    def calc_atomic_potentials(self, potential_range=8, oversample=2, kirkland=True):
        """Per-element projected-potential tables on the GPU (msa.py:137-146)."""
        if not kirkland:
            raise NotImplementedError("only the Kirkland parameterisation exists")
        torch = _torch()
        dev = self._dev()
        self.scattering_params = kirkland_params
        self.potential_range = float(potential_range)
        self.oversample = int(oversample)
        self.table_shape = tuple(int(np.floor(potential_range / p)) for p in self.pix_size)
        self.unique_Z = np.unique(self.supercell_Z)
        params = np.ascontiguousarray(kirkland_params[self.unique_Z - 1], dtype=np.float64)
        params_d = torch.from_numpy(params).to(dev)
        Py, Px = self.table_shape
        self.tables_device = torch.empty((len(self.unique_Z), Py, Px), dtype=torch.float32, device=dev)
        call("namsa_atomic_tables", ptr(params_d), len(self.unique_Z), Py, Px,
             self.potential_range, self.oversample, ptr(self.tables_device), stream_handle())
        host = self.tables_device.cpu().numpy()
        self.cached_pots = {int(Z): host[i] for i, Z in enumerate(self.unique_Z)}

    def site_table(self, slice_thickness):
        """Sites grouped by slice under the MSAGPU rule (msa.py:584-604).

        Returns (sites int32 [A][3] = (table index, int(y/py), int(x/px)),
        offsets int32 [S+1]).  Slice n holds float32(n dz) <= z < float32((n+1) dz);
        sites keep their input order inside a slice.
        """
        S = int(np.floor(np.float64(self.dims[-1]) / slice_thickness))
        z = self.supercell_xyz[:, 2]
        bounds = np.array([n * slice_thickness for n in range(S + 1)], dtype=np.float64).astype(np.float32)
        sl = np.searchsorted(bounds, z, side="right") - 1
        keep = np.nonzero((sl >= 0) & (sl < S))[0]
        order = keep[np.argsort(sl[keep], kind="stable")]
        zi = np.searchsorted(self.unique_Z, self.supercell_Z[order])
        ypix = (self.supercell_xyz[order, 1].astype(np.float64) / self.pix_size[0]).astype(np.int32)
        xpix = (self.supercell_xyz[order, 0].astype(np.float64) / self.pix_size[1]).astype(np.int32)
        sites = np.ascontiguousarray(np.stack([zi.astype(np.int32), ypix, xpix], axis=1))
        counts = np.bincount(sl[order], minlength=S) if len(order) else np.zeros(S, np.int64)
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        return sites, offsets

    def build_potential_slices(self, slice_thickness):
        """Transmission slices exp(i sigma V) on the GPU (msa.py:580-647)."""
        torch = _torch()
        dev = self._dev()
        if not hasattr(self, "tables_device"):
            self.calc_atomic_potentials()
        self.slice_t = float(slice_thickness)
        sites, offsets = self.site_table(slice_thickness)
        self.num_slices = np.int32(len(offsets) - 1)
        Ny, Nx = self._grid
        Py, Px = self.table_shape
        sites_d = torch.from_numpy(sites.reshape(-1)).to(dev) if len(sites) else None
        offs_d = torch.from_numpy(offsets).to(dev)
        self._t = torch.empty((int(self.num_slices), Ny, Nx), dtype=torch.complex64, device=dev)
        call("namsa_build_slices", ptr(self.tables_device), len(self.unique_Z), Py, Px,
             ptr(sites_d), ptr(offs_d), int(self.num_slices), Ny, Nx, np.float32(self.sigma),
             ptr(self._t), stream_handle())
        self.pot_dev_ptr = self._t
        self._slices_host = None
        self.print_verbose("Built %d potential slices with shape:%s pixels" % (self.num_slices, (Ny, Nx)))

    def integrate_potential_slices(self, pot, grid, slice_thickness, output=False):
        """Transmission slices from a sampled 3D potential, e.g. an all-electron
        ab initio one (msa.py:373-424 and :568-578).

        pot: [z][y][x] samples; grid: (z, y, x) coordinate arrays in Angstrom.
        As in the reference, the plane is cropped to the square min(Nx, Ny),
        S = int(z extent // slice_thickness), d = Z // S samples per slice
        (z cropped to S*d when d differs from ceil(Z/S)), and slice s is the
        negated trapezoid integral over its own d samples (no interval spans
        two slices); t = exp(i sigma V) is formed on the GPU
        (namsa_integrate_slices, FP64 integration).  Updates num_slices, dims
        (coordinate maxima of the full grids), sampling, kmax, max_ang,
        pix_size and kpix_size as msa.py:409-421 does.  output=True returns
        V [S][M][M] float64.
        """
        torch = _torch()
        dev = self._dev()
        z, y, x = (np.asarray(c, dtype=np.float64) for c in grid)
        pot = np.asarray(pot)
        self.slice_t = float(slice_thickness)
        S = int((z[-1] - z[0]) // slice_thickness)
        if S < 1:
            raise ValueError(f"z extent {z[-1] - z[0]} holds no slice of {slice_thickness}")
        if pot.shape[0] != z.size:
            raise ValueError(f"potential has {pot.shape[0]} z samples, the grid {z.size}")
        M = min(x.size, y.size)
        d = int(pot.shape[0] // S)
        pot_d = torch.from_numpy(np.ascontiguousarray(pot[: d * S, :M, :M], dtype=np.float64)).to(dev)
        z_d = torch.from_numpy(np.ascontiguousarray(z[: d * S])).to(dev)
        self._t = torch.empty((S, M, M), dtype=torch.complex64, device=dev)
        v_d = torch.empty((S, M, M), dtype=torch.float64, device=dev) if output else None
        call("namsa_integrate_slices", ptr(pot_d), ptr(z_d), S, d, M, M, float(self.sigma), ptr(self._t), ptr(v_d),
             stream_handle())
        del pot_d, z_d
        self.pot_dev_ptr = self._t
        self._slices_host = None
        self.num_slices = np.int32(S)
        self.dims = np.array([x.max(), y.max(), z.max()])
        if self.max_angle is None:
            self.sampling = np.array([M, M], dtype=np.int64)
            self.kmax = float(np.min(self.sampling / self.dims[:2]))
            self.max_ang = self.kmax * self.Lambda
        else:
            self.max_ang = self.max_angle
            self.kmax = self.max_ang / self.Lambda
            self.sampling = np.floor(self.kmax * self.dims[:2]).astype(np.int64)[::-1]
        self.pix_size = self.dims[:2][::-1] / self.sampling
        self.kpix_size = self.kmax / self.sampling
        if output:
            return v_d.cpu().numpy()
        return None

    @property
    def potential_slices(self):
        """Host copy of the complex64 transmission slices [S][Ny][Nx]."""
        if self._t is None:
            return None
        if self._slices_host is None:
            self._slices_host = self._t.cpu().numpy()
        return self._slices_host

    @property
    def potential_slices_device(self):
        return self._t

    # ------------------------------------------------------------------ probe
    def _probe_args(self, probe_dict):
        aber = probe_dict.get("aberration_dict", {"C1": 0.0, "C3": 0.0, "C5": 0.0})
        C1, C3, C5 = (np.float32(aber.get(k, 0.0)) for k in ("C1", "C3", "C5"))
        if probe_dict.get("scherzer", False):  # GPU rule: always replaces C1 (msa.py:677-679)
            C1 = np.float32(scherzer_defocus(self.Lambda, aber.get("C3", 0.0)))
        if not probe_dict.get("spherical_phase", True):
            C1 = C3 = C5 = np.float32(0.0)
        return C1, C3, C5

    def build_probe(self, probe_dict=DEFAULT_PROBE):
        """Probe psi_k, aperture, real-space psi and normalization (msa.py:669-735)."""
        torch = _torch()
        dev = self._dev()
        self._check_grid()
        self.probe_dict = dict(probe_dict)
        Ny, Nx = self._grid
        C1, C3, C5 = self._probe_args(probe_dict)
        smooth = bool(probe_dict.get("smooth_apert", True))
        steep = float(probe_dict.get("gpu_apert_smooth", GPU_APERTURE_STEEPNESS))
        self.psi_k_device = torch.empty((Ny, Nx), dtype=torch.complex64, device=dev)
        apert = torch.empty((Ny, Nx), dtype=torch.float32, device=dev)
        st = stream_handle()
        call("namsa_probe_k", Ny, Nx, np.float32(self.kmax), np.float32(self.semi_ang / self.Lambda),
             np.float32(self.Lambda), C1, C3, C5, int(smooth), np.float32(steep),
             ptr(self.psi_k_device), ptr(apert), st)
        # incident probe at the origin = fftshift(IFFT_u(psi_k)), and its norm
        pos0 = torch.zeros((1, 2), dtype=torch.float32, device=dev)
        psi = torch.empty((1, Ny, Nx), dtype=torch.complex64, device=dev)
        call("namsa_multislice", ptr(self.psi_k_device), ptr(pos0), 1, Ny, Nx, np.float32(self.kmax),
             None, 0, None, 0, Nx, 0, _lib.OUT_INCIDENT, np.float32(1.0), ptr(psi), None, None, 0,
             None, None, st)
        norm = torch.empty(1, dtype=torch.float64, device=dev)
        call("namsa_norm2", ptr(psi), Ny * Nx, ptr(norm), st)
        self.normalization = float(norm.item())
        self.psi_k = self.psi_k_device.cpu().numpy()
        self.apert = apert.cpu().numpy()
        self.psi = (psi[0].cpu().numpy() / np.sqrt(self.normalization)).astype(np.complex64)

    # ------------------------------------------------------------------ raster
    def generate_probe_positions(self, probe_step=np.array([0.1, 0.1]),
                                 probe_range=np.array([[0.0, 1.0], [0.0, 1.0]]), grid_steps=None,
                                 fraction=0.5, origin=np.array([0, 0])):
        """Probe raster [P][2] = (y, x) in Angstrom (msa.py:253-278)."""
        d = self.dims[:2].astype(np.float64)
        if grid_steps is not None:
            gs = np.asarray(grid_steps).astype(np.int64)
            org = np.asarray(origin, dtype=np.float64)[::-1]
            start = (0.5 + org * d - d * fraction) / 2
            stop = (0.5 + org * d + d * fraction) / 2
            xs = np.linspace(start[0], stop[0], gs[0])
            ys = np.linspace(start[1], stop[1], gs[1])
            rx, ry = [np.array([start[i], stop[i]]) for i in range(2)]
        else:
            pr = np.asarray(probe_range, dtype=np.float64)
            gs = np.floor(np.diff(pr).flatten() * d / np.asarray(probe_step)).astype(np.int64)
            rx, ry = [(pr[i] - 0.5) * d[i] for i in range(2)]
            xs = np.linspace(rx[0], rx[1], gs[0])
            ys = np.linspace(ry[0], ry[1], gs[1])
        # x is the slow axis, y the fast one; each row is (y, x)
        self.probe_positions = np.stack([np.tile(ys, len(xs)), np.repeat(xs, len(ys))], axis=1)
        self.grid_steps = np.array([gs[0], gs[1]])
        self.grid_range = np.array([rx, ry]).flatten()
        self.num_probes = np.int32(self.probe_positions.shape[0])

    def plan_simulation(self, num_probes=None, group=None):
        """Largest number of probe positions whose device buffers fit half the
        free memory (msa.py:445-461).

        Per position: 16 N^2 bytes, the worst case of the output (two
        double-buffered complex64 chunk planes when the run goes in phases);
        plus one probe group of wave state and the resident slices.  As in the
        reference, the count shrinks by 3/4 until that fits.  ``multislice``
        runs more positions than this in phases (``phase_plan``).
        """
        torch = _torch()
        self._dev()
        if num_probes is None:
            num_probes = int(self.num_probes)
        free, _ = torch.cuda.mem_get_info(self.device)
        Ny, Nx = self._grid
        plane = Ny * Nx
        slices = 0 if self._t is None else int(self._t.numel()) * 8
        G = int(group) if group else cache_group(Ny, Nx)

        def need(n):
            return n * 16 * plane + min(n, G) * 8 * plane + slices

        n = int(num_probes)
        while n > 1 and need(n) > 0.5 * free:
            n = int(n * 0.75)
        self.max_probes = max(1, n)
        self.print_verbose("Device can hold at most %d probes simultaneously." % self.max_probes)
        return self.max_probes

    # ------------------------------------------------------------------ propagation
    def build_propagator(self, bandwidth=1 / 3):
        """Propagator * bandwidth mask on the device (msa.py:833-839)."""
        torch = _torch()
        dev = self._dev()
        Ny, Nx = self._grid
        self.bandwidth = bandwidth
        self.pm_device = torch.empty((Ny, Nx), dtype=torch.complex64, device=dev)
        self.mask_device = torch.empty((Ny, Nx), dtype=torch.float32, device=dev)
        call("namsa_propagator", Ny, Nx, np.float32(self.kmax), np.float32(bandwidth * self.kmax),
             np.float32(self.Lambda), np.float32(getattr(self, "slice_t", 0.0)),
             ptr(self.pm_device), ptr(self.mask_device), stream_handle())
        # columns the bandwidth mask leaves open: outside them every slice's
        # spectrum is exactly zero and the kernels skip the work
        cols = np.nonzero(self.mask_device.cpu().numpy().any(axis=0))[0]
        self.band = (int(cols[0]), int(cols[-1]) + 1) if len(cols) else (0, Nx)
        # where the CBED slice loop keeps the wave in 2 x 4 tiles, a tiled
        # copy of the propagator lets the column pass read it the same way
        self.pm_tiled_device = None
        if _lib.load().namsa_wave_tiled(Ny, Nx):
            self.pm_tiled_device = torch.empty_like(self.pm_device)
            call("namsa_tile_plane", ptr(self.pm_device), ptr(self.pm_tiled_device), Ny, Nx, stream_handle())
        return self.pm_device

    def cbed_propagator(self):
        """(propagator pointer, mode flags) for namsa_multislice in CBED or
        exit-wave mode: the tiled copy with NAMSA_PM_TILED where the slice
        loop tiles the wave, else the row-major propagator."""
        if getattr(self, "pm_tiled_device", None) is not None:
            return ptr(self.pm_tiled_device), _lib.PM_TILED
        return ptr(self.pm_device), 0

    def multislice(self, bandwidth=1 / 3, unified_mem=False, batch_size=None, transmit=True,
                   detectors=None, output="auto", group=None, moments=False, sink=None, records=False):
        """Propagate every probe position through all slices (msa.py:767-924).

        transmit=True: ``probes`` = float32 CBED [P][Ny][Nx], normalised by the
        probe's normalization (sum <= 1).  transmit=False: complex incident
        probes / sqrt(normalization).  output="exit_wave" returns real-space
        exit waves / sqrt(normalization) instead.  ``detectors`` ([nDet][Ny][Nx]
        float weights) fuses the detector integrals into the readout; the
        result lands in ``self.detector_signals`` [P][nDet] (float64).
        ``group`` probes go through all slices together (default: as many as
        fit the last-level cache, see ``cache_group``).  moments=True (CBED
        output) also keeps each CBED's sum and sum of squares, taken in the
        readout, in ``cbed_moments_device`` [P][2] float64, for
        ``records.process_cbed(..., moments=...)``.

        Memory: when every position fits at once (``plan_simulation``'s
        ``max_probes``) and there is no sink, the output stays on the device
        (``probes_device``; ``probes`` copies it to host on first use).
        Otherwise the positions go in phases of at most ``max_probes``
        (``batch_size`` if smaller), in order, as the reference's phases do
        (msa.py:843-907): the device holds two chunks of output, and each
        finished chunk is copied on a second stream to pinned host memory
        while the next one computes -- into ``probes`` (a pinned [P][Ny][Nx]
        host array), or to a sink.  ``probes_device`` is then None.

        sink(start, array): stream the CBEDs out while the simulation runs.
        The probes go in chunks (``batch_size``, by default a multiple of
        the group giving about eight chunks). Each chunk's CBEDs, or with
        records=True its float16 records (encoded on the GPU from the
        readout's moments), are copied to pinned host memory on a second
        stream, while the next chunk computes. sink then gets the chunk's
        first probe index and a host array, which is only valid during the
        call. Nothing else keeps the CBEDs: ``probes`` and ``probes_device``
        are None.
        """
        torch = _torch()
        dev = self._dev()
        self._check_grid()
        if self._t is None and transmit:
            warn("Potential slices must be calculated first before calling multi_slice.")
            return None
        if not hasattr(self, "psi_k_device"):
            warn("Probe in k-space must be initiated first before calling multi_slice.")
            return None
        if getattr(self, "probe_positions", None) is None:
            warn("Probe positions must be calculated first before calling multislice.")
            return None
        Ny, Nx = self._grid
        P = int(self.num_probes)
        group = int(group) if group else cache_group(Ny, Nx)
        if not hasattr(self, "max_probes"):
            self.plan_simulation(group=group)
        S = int(self.num_slices) if transmit else 0
        self.build_propagator(bandwidth)
        mode = _lib.OUT_CBED if transmit else _lib.OUT_INCIDENT
        if output == "exit_wave":
            mode = _lib.OUT_EXIT
            S = int(self.num_slices)
        if sink is not None and mode != _lib.OUT_CBED:
            raise NamsaError("sink streams CBEDs: it needs transmit=True and CBED output")
        chunk, resident = phase_plan(P, self.max_probes, group, sink is not None, batch_size)
        ws_planes = min(chunk, group)
        pos_all = torch.from_numpy(np.ascontiguousarray(self.probe_positions, dtype=np.float32)).to(dev)
        if mode == _lib.OUT_CBED:
            odt = torch.float32
            psi_ws = torch.empty((ws_planes, Ny, Nx), dtype=torch.complex64, device=dev)
            scale = np.float32(1.0 / self.normalization)
        else:   # the wave state itself is the output
            odt = torch.complex64
            scale = np.float32(1.0 / np.sqrt(self.normalization))
        det = None
        nDet = 0
        if detectors is not None and mode == _lib.OUT_CBED:
            det = torch.as_tensor(np.asarray(detectors, dtype=np.float32).reshape(-1, Ny, Nx)).to(dev)
            nDet = det.shape[0]
            lib = _lib.load()
            det_ws = torch.empty(int(lib.namsa_det_workspace_doubles(chunk, Ny, Nx, nDet)),
                                 dtype=torch.float64, device=dev)
            self._det_dev = torch.empty((P, nDet), dtype=torch.float64, device=dev)
        self._moments_dev = None
        with_moments = (bool(moments) or (sink is not None and records)) and mode == _lib.OUT_CBED
        if with_moments:
            mom_ws = torch.empty(int(_lib.load().namsa_det_workspace_doubles(chunk, Ny, Nx, 2)),
                                 dtype=torch.float64, device=dev)
            self._moments_dev = torch.empty((P, 2), dtype=torch.float64, device=dev)
        st = stream_handle()
        t_arg = ptr(self._t) if S > 0 else None
        self._probes_dev = None
        self._probes_host = None
        self._host_pinned = None
        if resident:
            out = torch.empty((P, Ny, Nx), dtype=odt, device=dev)
        else:
            slots = [torch.empty((chunk, Ny, Nx), dtype=odt, device=dev) for _ in range(2 if P > chunk else 1)]
            if sink is None:
                self._host_pinned = torch.empty((P, Ny, Nx), dtype=odt, pin_memory=True)
            streamer = _ChunkStreamer(Ny, Nx, odt, dev, chunk, len(slots), sink=sink, host_out=self._host_pinned,
                                      moments=self._moments_dev if (sink is not None and records) else None)
        for k, start in enumerate(range(0, P, chunk)):
            B = min(chunk, P - start)
            if resident:
                dst = out[start:start + B]
            else:
                slot = k % len(slots)
                streamer.claim(slot)
                dst = slots[slot][:B]
            if mode == _lib.OUT_CBED:
                cbed_ptr, ws_ptr = ptr(dst), ptr(psi_ws)
            else:
                cbed_ptr, ws_ptr = None, ptr(dst)
            det_args = (ptr(det), nDet, ptr(det_ws), ptr(self._det_dev[start])) if nDet else (None, 0, None, None)
            if with_moments:
                call("namsa_multislice_moments", ptr(self.psi_k_device), ptr(pos_all[start]), B, Ny, Nx,
                     np.float32(self.kmax), t_arg, S, ptr(self.pm_device), self.band[0], self.band[1],
                     group, scale, ws_ptr, cbed_ptr, *det_args, ptr(mom_ws), ptr(self._moments_dev[start]), st)
            else:
                pm_arg, flags = (self.cbed_propagator() if mode in (_lib.OUT_CBED, _lib.OUT_EXIT)
                                 else (ptr(self.pm_device), 0))
                call("namsa_multislice", ptr(self.psi_k_device), ptr(pos_all[start]), B, Ny, Nx,
                     np.float32(self.kmax), t_arg, S, pm_arg, self.band[0], self.band[1],
                     group, mode | flags, scale, ws_ptr, cbed_ptr, *det_args, st)
            if not resident:
                streamer.push(dst, start, B, slot)
        if resident:
            self._probes_dev = out
        else:
            streamer.finish()
            if self._host_pinned is not None:
                self._probes_host = self._host_pinned.numpy()
        self.detector_signals = self._det_dev.cpu().numpy() if nDet else None
        self.vars.append(self._t)
        return None

    @property
    def probes(self):
        """Host copy of the multislice output (CBED float32 or complex probes);
        None after a run that streamed its output to a sink."""
        if self._probes_host is not None:
            return self._probes_host
        if self._probes_dev is None:
            return None
        if self._probes_host is None:
            self._probes_host = self._probes_dev.cpu().numpy()
        return self._probes_host

    @property
    def probes_device(self):
        return self._probes_dev

    @property
    def cbed_moments_device(self):
        """[P][2] float64 (sum, sum of squares) of each CBED from multislice(moments=True), else None."""
        return getattr(self, "_moments_dev", None)

    # ------------------------------------------------------------------ detectors
    def detector_mask(self, inner_angle, outer_angle, mode="reference"):
        """Annular detector weights [Ny][Nx] float32.

        mode="reference": integrate_cbed's pixel-radius annulus (msa.py:426-435),
        radius = N * angle / (2 * max_ang) on the integer grid centred at N/2.
        mode="physical": inner <= k*lambda < outer on the linspace k-grid the
        CBED is sampled on.
        """
        Ny, Nx = self._grid
        if mode == "reference":
            if not (inner_angle < outer_angle < self.max_ang):
                raise ValueError("Detector angles exceed maximum scattering angle simulated "
                                 "and/or values are not consistent")
            g0, g1 = np.meshgrid(np.arange(-Ny // 2, Ny // 2), np.arange(-Nx // 2, Nx // 2), indexing="ij")
            r = np.sqrt(g0 ** 2 + g1 ** 2)
            n = max(Ny, Nx)
            inner = n * (inner_angle / self.max_ang / 2) - r > 0
            outer = n * (outer_angle / self.max_ang / 2) - r > 0
            return np.logical_xor(inner, outer).astype(np.float32)
        if mode == "physical":
            km = np.float64(np.float32(self.kmax))
            ky = np.arange(Ny) * km / (Ny - 1) - km / 2
            kx = np.arange(Nx) * km / (Nx - 1) - km / 2
            theta = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2) * self.Lambda
            return ((theta >= inner_angle) & (theta < outer_angle)).astype(np.float32)
        raise ValueError("mode must be 'reference' or 'physical'")

    def _cbed_blocks(self):
        """(start, device float32 CBEDs) blocks of the last run's output: the
        resident device output in one block, or a phased run's pinned host
        output uploaded one chunk at a time."""
        torch = _torch()
        if self._probes_dev is not None:
            if self._probes_dev.dtype != torch.float32:
                raise NamsaError("integrate_cbed needs CBEDs from multislice(transmit=True)")
            yield 0, self._probes_dev
            return
        host = getattr(self, "_host_pinned", None)
        if host is None or host.dtype != torch.float32:
            raise NamsaError("integrate_cbed needs CBEDs from multislice(transmit=True) kept in probes "
                             "(a run that streamed them to a sink keeps none)")
        step = max(1, min(int(getattr(self, "max_probes", host.shape[0])), host.shape[0]))
        for start in range(0, host.shape[0], step):
            yield start, host[start:start + step].to(self.device, non_blocking=True)

    def integrate_cbed(self, detector_array=None,
                       detector_params={"inner_angle": 50e-3, "outer_angle": 100e-3}):
        """Detector signal per raster point, shaped grid_steps (msa.py:426-439)."""
        torch = _torch()
        dev = self._dev()
        if detector_array is None:
            detector_array = self.detector_mask(detector_params["inner_angle"],
                                                detector_params["outer_angle"])
        Ny, Nx = self._grid
        det = torch.as_tensor(np.asarray(detector_array, dtype=np.float32).reshape(1, Ny, Nx)).to(dev)
        out = torch.empty((int(self.num_probes), 1), dtype=torch.float64, device=dev)
        for start, cb in self._cbed_blocks():
            call("namsa_integrate", ptr(cb), cb.shape[0], Ny, Nx, ptr(det), 1, ptr(out[start]), stream_handle())
        signal = out[:, 0].cpu().numpy().reshape([int(g) for g in self.grid_steps])
        return signal, np.asarray(detector_array)

    def check_simulation(self):
        """Integrated intensity per probe; warns outside [0.95, 1.0] (msa.py:363-371)."""
        torch = _torch()
        Ny, Nx = self._grid
        st = stream_handle()
        cb = self._probes_dev
        if cb is not None and cb.dtype != torch.float32:
            out = torch.empty((cb.shape[0],), dtype=torch.float64, device=cb.device)
            for i in range(cb.shape[0]):
                call("namsa_norm2", ptr(cb[i]), Ny * Nx, ptr(out[i]), st)
        elif cb is not None or getattr(self, "_host_pinned", None) is not None:
            if cb is None and self._host_pinned.dtype != torch.float32:
                raise NamsaError("check_simulation of phased complex outputs is not supported")
            out = torch.empty((int(self.num_probes),), dtype=torch.float64, device=self.device)
            ones = torch.ones((1, Ny, Nx), dtype=torch.float32, device=self.device)
            for start, blk in self._cbed_blocks():
                call("namsa_integrate", ptr(blk), blk.shape[0], Ny, Nx, ptr(ones), 1, ptr(out[start]), st)
        else:
            return None
        prob = out.cpu().numpy()
        print("Max (Min) Integrated Intensity: %2.2f (%2.2f)" % (prob.max(), prob.min()))
        if prob.max() > 1.0 or prob.min() < 0.95:
            print("Significant deviation of the probability from unity is found.\n"
                  "Change the sampling and/or slice thickness!")
        return prob
