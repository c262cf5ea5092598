"""Training-record encoders and containers (namsa_amd.records) against the
oracle (oracle/records_oracle.py), whose functions are pinned here by worked
values, invariants and an explicit convolution. CPU only; the GPU kernels
are compared in tests/test_gpu_records.py."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from namsa_amd import records as REC
from oracle import records_oracle as R


def test_process_cbed_worked_values():
    # x = [0, 1, 2, 3]: mean 1.5, population deviation sqrt(1.25)
    got = R.process_cbed(np.array([[[0.0, 1.0], [2.0, 3.0]]]))
    want = (np.array([0.0, 1.0, 2.0, 3.0]) - 1.5) / np.sqrt(1.25)
    assert got.dtype == np.float16 and got.shape == (1, 2, 2)
    np.testing.assert_array_equal(got.ravel(), want.astype(np.float16))


def test_process_cbed_invariants():
    rng = np.random.default_rng(0)
    x = rng.exponential(size=(5, 32, 48)).astype(np.float32) ** 3
    y = R.process_cbed(x).astype(np.float64)
    np.testing.assert_allclose(y.mean(axis=(1, 2)), 0, atol=2e-3)
    np.testing.assert_allclose(y.std(axis=(1, 2)), 1, atol=2e-3)
    z = R.process_cbed(3.5 * x + 7.0).astype(np.float64)          # affine invariance, to half rounding
    assert np.abs(z - y).max() <= 2 * 2.0 ** -10 * max(1.0, np.abs(y).max())
    c = R.process_cbed(np.full((2, 4, 4), 0.25, np.float32))       # constant item: 0 / 0
    assert np.isnan(c).all()


def test_gaussian_filter_restatement():
    """scipy's gaussian_filter (the reference's call) equals an explicit
    separable convolution with symmetric padding, also when the radius
    exceeds the image (repeated reflections)."""
    rng = np.random.default_rng(1)
    for shape in [(40, 37), (6, 5), (64, 64)]:
        img = rng.normal(size=shape).astype(np.float32)
        np.testing.assert_array_equal(gaussian_filter(img, 1.2), R.gaussian_direct(img, 1.2))


def test_process_potential_shape_and_floor():
    rng = np.random.default_rng(2)
    t = np.exp(1j * rng.uniform(-0.3, 0.3, (3, 256, 256))).astype(np.complex64)
    rec = R.process_potential(t)
    assert rec.shape == (1, 128, 128) and rec.dtype == np.float16        # sim_prod.py:127's expected shape
    assert R.crop_slice((256, 256)) == slice(64, 192)
    flat = R.process_potential(np.ones((2, 64, 64), np.complex64))       # imag 0 everywhere: 0 / floor = 0
    assert flat.shape == (1, 32, 32) and (flat == 0).all()
    r = R.process_potential(t[:, :200, :256])                             # rectangular: slice(50, 192) on both axes
    assert r.shape == (1, 142, 142)


def test_lmdb_headers_and_records():
    h = REC.lmdb_header_entries()
    assert h == R.lmdb_headers()
    d = dict(h)
    assert np.frombuffer(d[b"input_shape"], dtype=np.int64).tolist() == [64, 256, 256]
    assert np.frombuffer(d[b"output_shape"], dtype=np.int64).tolist() == [1, 128, 128]
    assert d[b"output_name"] == b"2d_potential_" and d[b"input_name"] == b"cbed_"
    assert d[b"header_entries"] == b"\x00" * 6 and [k for k, _ in h][-1] == b"header_entries"

    class Txn(dict):
        def put(self, k, v):
            self[k] = v

    txn = Txn()
    REC.write_lmdb_headers(txn)
    cb = np.arange(2 * 4 * 4, dtype=np.float16).reshape(2, 4, 4)
    pot = np.ones((1, 2, 2), np.float16)
    REC.write_lmdb(txn, 7, cb, pot)
    assert txn[b"cbed_7"] == cb.tobytes() and txn[b"2d_potential_7"] == pot.tobytes()
    assert len(txn) == 9


class _Attrs(dict):
    pass


class _Group:
    def __init__(self):
        self.children, self.attrs = {}, _Attrs()

    def items(self):
        return list(self.children.items())

    def create_group(self, name):
        self.children[name] = g = _Group()
        return g

    def create_dataset(self, name, shape=None, dtype=None, data=None):
        self.children[name] = arr = np.array(data) if data is not None else np.zeros(shape, dtype=dtype)
        return arr

    def __contains__(self, name):
        return name in self.children

    def __getitem__(self, name):
        return self.children[name]


def test_h5_layouts_through_a_group_interface():
    root = _Group()
    params = {"energy": 100.0, "probe_params": {"smooth_apert": True,
                                                "aberration_dict": {"C1": 0.0, "C3": 1e3, "C5": 0.0}}}
    g = REC.write_h5(root, np.zeros((2, 4, 4), np.float16), np.ones((1, 2, 2), np.float16), params)
    assert list(root.children) == ["sample_0"] and set(g.children) == {"CBED", "potential"}
    assert g.attrs["energy"] == 100.0 and g.attrs["smooth_apert"] is True
    np.testing.assert_array_equal(g.attrs["aberrations"], [0.0, 1e3, 0.0])
    REC.write_h5(root, np.zeros((2, 4, 4)), np.zeros((1, 2, 2)), {})
    assert list(root.children) == ["sample_0", "sample_1"]
    f = _Group()                                                 # MSAMPI '4D_CBED': rows at each part's offset
    REC.write_4d_cbed(f, np.ones((3, 4, 4), np.float32), total_num_probes=5, offset=2)
    REC.write_4d_cbed(f, np.full((2, 4, 4), 2, np.float32), total_num_probes=5, offset=0)
    np.testing.assert_array_equal(f["4D_CBED"][:, 0, 0], [2, 2, 1, 1, 1])
