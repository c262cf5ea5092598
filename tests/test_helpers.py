"""Host-side helpers with counterparts in the reference's namsa/optics.py and
namsa/utils.py, against closed forms and explicit loops."""
import numpy as np

from namsa_amd import optics as OP
from namsa_amd import utils as U
from namsa_amd.supercell import lattice_matrix


def test_dhkl_spacing_closed_forms():
    a = 5.43
    cubic = np.eye(3) * a
    assert np.isclose(U.dhkl_spacing(a ** 3, cubic, [0, 0, 1]), a)
    assert np.isclose(U.dhkl_spacing(a ** 3, cubic, [1, 1, 0]), a / np.sqrt(2))
    assert np.isclose(U.dhkl_spacing(a ** 3, cubic, [1, 1, 1]), a / np.sqrt(3))
    ah, c = 3.189, 5.185
    hexa = lattice_matrix(ah, ah, c, 90, 90, 120)
    vol = abs(np.linalg.det(hexa))
    assert np.isclose(U.dhkl_spacing(vol, hexa, [1, 0, 0]), ah * np.sqrt(3) / 2)
    assert np.isclose(U.dhkl_spacing(vol, hexa, [0, 0, 2]), c / 2)


def test_bin_2d_array_against_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(6, 8))
    want = np.array([[x[2 * i:2 * i + 2, 2 * j:2 * j + 2].mean() for j in range(4)] for i in range(3)])
    np.testing.assert_allclose(U.bin_2d_array(x, 2), want)
    np.testing.assert_allclose(U.bin_2d_array(x, 2, mode="sum"), 4 * want)


def test_image_tile_layout():
    rng = np.random.default_rng(1)
    d = rng.normal(size=(5, 3, 2))
    t = U.imageTile(d, padsize=1, padval=-9.0)
    assert t.shape == (3 * 4, 3 * 3)
    for k in range(9):
        r, c = (k // 3) * 4, (k % 3) * 3
        block = t[r:r + 3, c:c + 2]
        np.testing.assert_array_equal(block, d[k] if k < 5 else np.full((3, 2), -9.0))
        assert (t[r + 3, c:c + 3] == -9.0).all() and (t[r:r + 4, c + 2] == -9.0).all()


def test_optics_closed_forms():
    lam = OP.voltage2Lambda(200e3)
    assert np.isclose(OP.transverse_width(lam, 1e-3), lam / (2 * np.pi * 1e-3))
    for cs in (0.5, 1.0, 2.0):                     # the product is independent of Cs
        prod = OP.semi_angle_limit(lam, cs) * OP.no_overlap_dhkl(lam, cs)
        assert np.isclose(prod, 0.7874 * 1.51 * lam * 1e3)
    theta_mrad, cs_ret, defocus = OP.overlap_params(0.5, 2.0, lam)
    assert np.isclose(theta_mrad, 0.5 * lam / 4.0 * 1e3)
    # C3 = (theta/1.51)^-4 lambda in Angstrom, returned x1e-6 (ten times mm): the Scherzer limit of
    # that Cs, read as mm, is theta * 10^(-1/4)
    assert np.isclose(OP.semi_angle_limit(lam, cs_ret), theta_mrad * 10 ** -0.25)
    c3 = cs_ret * 1e6
    assert np.isclose(defocus, -1.15 * c3 ** -0.25 * lam ** -0.75)
    assert np.isclose(OP.cal_probe_size(lam, 0.02), 0.61 * lam / 0.02) and OP.cal_probe_size(lam, 0.02, True) is None


def test_spherical_phase_error():
    lam = OP.voltage2Lambda(100e3)
    k = np.array([0.0, 0.5, 1.0])
    got = OP.spherical_phase_error(k, lam, False, C1=50.0, C3=1e4, C5=0.0)
    q = k * lam
    chi = 2 * np.pi / lam * (-25.0 * q ** 2 + 2500.0 * q ** 4)
    np.testing.assert_allclose(got, np.exp(-1j * chi))
    sch = OP.spherical_phase_error(k, lam, True, C1=0.0, C3=1e4, C5=0.0)
    c1 = np.sqrt(1.5) * np.sqrt(1e4 * lam)
    np.testing.assert_allclose(sch, np.exp(-1j * 2 * np.pi / lam * (-0.5 * c1 * q ** 2 + 2500.0 * q ** 4)))
