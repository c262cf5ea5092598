"""Structure front-end (namsa_amd.supercell) and its oracle
(oracle/supercell_oracle.py) against the reference's own fixtures
(cif_files/, copied under tests/golden/cif_files; the reference's tests are
tests/supercell_tests.py:8-54), closed forms, and each other. CPU only."""
import os

import numpy as np
import pytest

from namsa_amd import supercell as S
from oracle import supercell_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "cif_files")
MVC = os.path.join(GOLD, "mvc-9418.cif")
SI = os.path.join(GOLD, "Si.cif")
XYZ = [("atom_type", "|U16"), ("x", "f"), ("y", "f"), ("z", "f")]


def _fixture(name):
    return np.loadtxt(os.path.join(GOLD, name), skiprows=2, dtype=XYZ)


def _pos(a):
    return np.column_stack([a["x"], a["y"], a["z"]]).astype(np.float64)


# files hold 4 decimals (5e-5) and the records are float32 (a few 1e-6 at 60 Å)
FILE_TOL = 6e-5


def _assert_same_cell(got, ref, tol=FILE_TOL):
    assert len(got) == len(ref)
    assert (got["atom_type"] == ref["atom_type"]).all()
    assert np.abs(_pos(got) - _pos(ref)).max() < tol


# --- the reference's three fixture tests, then stronger forms -------------------------------------------------


def test_unit_cell_matches_reference_fixture():
    """supercell_tests.py:10-23: mvc-9418 unit cell (triclinic P1, 45
    atoms): positions equal at 2 decimals and in order; here also within the
    file's rounding. Pins the CIF reading and the pymatgen lattice
    convention."""
    sp = S.SupercellBuilder(MVC)
    sp.build_unit_cell()
    ref = _fixture("mvc-9418_P_1_1_1.xyz")
    assert (np.round(_pos(sp.xyz_positions), 2) == np.round(_pos(ref), 2)).all()
    _assert_same_cell(sp.xyz_positions, ref)


@pytest.mark.parametrize("scale,name", [(1, "mvc-9418_A_1_1_1.xyz"), (3, "mvc-9418_A_3_3_3.xyz")])
def test_centered_transform_matches_reference_fixture(scale, name):
    """supercell_tests.py:25-39 (A centering, 3x3x3: 2430 atoms; its check is
    a signed sum of differences) and the A 1x1x1 fixture, compared atom by
    atom. Positions reach 60 Å, where 6-decimal fractional rounding adds
    up to 4e-5."""
    sp = S.SupercellBuilder(MVC)
    sp.transform_unit_cell(centering_type="A", scaling=np.array([scale] * 3))
    sp.build_unit_cell()
    _assert_same_cell(sp.xyz_positions, _fixture(name), tol=1.2e-4)


def test_site_search_matches_p333_fixture():
    """mvc-9418_P_3_3_3.xyz (1215 atoms) is the site search run on the
    P·3 cell; the reference's P branch skips the search today, so the
    search is called directly, in the product and in the oracle."""
    sp = S.SupercellBuilder(MVC)
    sites0, ops0, basis0 = sp.sites.copy(), list(sp.sym_ops), sp.basis.copy()
    sp.transform_unit_cell(scaling=np.array([3, 3, 3]))
    assert sp.sites.size == 45                       # P: no search, as in the reference
    sp.sites = sp._search_sites(sp.sites, sp.Q_mat, np.array([3, 3, 3]))
    sp.build_unit_cell()
    ref = _fixture("mvc-9418_P_3_3_3.xyz")
    _assert_same_cell(sp.xyz_positions, ref, tol=1.2e-4)

    basis, _, _, ops, _ = O.transform(basis0, ops0, sites0, scaling=(3, 3, 3))
    found = O.search_sites(sites0, np.linalg.inv(np.diag([3.0, 3.0, 3.0, 1.0])), (3, 3, 3))
    _, cart, types = O.unit_cell(basis, ops, found)
    assert (types == ref["atom_type"]).all() and np.abs(cart - _pos(ref)).max() < 1.2e-4


def test_supercell_matches_reference_fixture_with_lattice_rule():
    """supercell_tests.py:41-55: default transform, 20 Å box. The fixture
    (726 atoms) is reproduced by the lattice-vector projection rule, and
    to_xyz writes it byte for byte."""
    sp = S.SupercellBuilder(MVC)
    sp.transform_unit_cell()
    sp.build_unit_cell()
    sp.make_orthogonal_supercell(projection_rule="lattice")
    ref = _fixture("supercell_mvc-9418.xyz")
    _assert_same_cell(sp.supercell_xyz_positions, ref)
    orc = O.orthogonal_supercell(sp.lattice_const, sp.lattice_angles, sp.unit_cell_positions, rule="lattice")
    _assert_same_cell(orc, ref)


def test_supercell_file_is_byte_identical(tmp_path):
    sp = S.SupercellBuilder(MVC)
    sp.transform_unit_cell()
    sp.build_unit_cell()
    sp.make_orthogonal_supercell(projection_rule="lattice", xyz=str(tmp_path / "t.xyz"),
                                 XYZ=str(tmp_path / "t.XYZ"))
    with open(os.path.join(GOLD, "supercell_mvc-9418.xyz")) as f:
        assert (tmp_path / "t.xyz").read_text() == f.read()


# --- the reference's element-wise projection (the default) ------------------------------------------------------


def test_components_rule_frame_closed_form():
    """Default projections [1,0,0] and [0,0,1] under the element-wise rule:
    the new axes are a = (0,-1,0), b = (1,0,0), c = (0,0,1), so trans_mat
    is that rotation times the normalised Cartesian basis."""
    sp = S.SupercellBuilder(MVC)
    sp.transform_unit_cell()
    T, t_vec, lo, hi, bnorm, size = sp._supercell_frame([20.0] * 3, [1, 0, 0], [0, 0, 1], "components")
    cart = sp.CartesianBasis()
    rot = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    np.testing.assert_allclose(T, rot @ (cart / np.linalg.norm(cart, axis=0)), atol=1e-15)
    np.testing.assert_allclose(t_vec, rot @ cart, atol=1e-12)      # translations = rotated lattice vectors
    np.testing.assert_allclose(bnorm, sp.lattice_const, rtol=1e-14)


def test_components_rule_matches_oracle():
    sp = S.SupercellBuilder(MVC)
    sp.transform_unit_cell()
    sp.build_unit_cell()
    sp.make_orthogonal_supercell(supercell_size=np.array([25.0, 18.0, 30.0]), projec_1=np.array([1, 1, 0]),
                                 projec_2=np.array([0, 0, 1]))
    orc = O.orthogonal_supercell(sp.lattice_const, sp.lattice_angles, sp.unit_cell_positions,
                                 size=(25.0, 18.0, 30.0), projec_1=(1, 1, 0), projec_2=(0, 0, 1))
    got = sp.supercell_sites
    assert len(got) == len(orc) > 1000
    for f in ("atom_type", "atomic_number", "occ", "DW"):
        assert (got[f] == orc[f]).all()
    assert np.abs(_pos(got) - _pos(orc)).max() < 1e-5
    pos = _pos(got)
    assert (pos >= 0).all() and (pos <= np.array([25.0, 18.0, 30.0]) + 1e-5).all()


# --- closed forms on silicon --------------------------------------------------------------------------------------


def _diamond(a):
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    return np.concatenate([fcc, fcc + 0.25]) * 1.0


def test_si_unit_cell_is_diamond():
    """192 Fd-3m operations (origin choice 2) on Si at the origin give the 8
    diamond positions."""
    sp = S.SupercellBuilder(SI)
    assert len(sp.sym_ops) == 192
    sp.build_unit_cell()
    frac = np.column_stack([sp.unit_cell_positions[k] for k in ("frac_x", "frac_y", "frac_z")])
    want = _diamond(5.43094)
    assert len(frac) == 8
    assert sorted(map(tuple, np.round(frac, 9))) == sorted(map(tuple, want))
    np.testing.assert_allclose(_pos(sp.xyz_positions), frac * 5.43094, atol=1e-5)
    _, cart, _ = O.unit_cell(sp.basis, [O.symop(s) for s in sp.cif_data["_space_group_symop_operation_xyz"]],
                             sp.sites)
    np.testing.assert_allclose(cart, _pos(sp.xyz_positions), atol=1e-5)


@pytest.mark.parametrize("rule", ["components", "lattice"])
def test_si_supercell_against_lattice_enumeration(rule):
    """Si in a 13 x 11 x 9.5 Å box with y along [110] and z along [001]: the
    atoms equal the diamond lattice points, rotated into the box frame, that
    fall inside it (enumerated directly; atoms within 1e-6 Å of a face may be
    on either side)."""
    a, size = 5.43094, np.array([13.0, 11.0, 9.5])
    sp = S.SupercellBuilder(SI)
    sp.build_unit_cell()
    sp.make_orthogonal_supercell(supercell_size=size, projec_1=np.array([1, 1, 0]), projec_2=np.array([0, 0, 1]),
                                 projection_rule=rule)
    got = _pos(sp.supercell_sites)
    y = np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    z = np.array([0.0, 0.0, 1.0])
    x = np.cross(y, z)
    n = np.arange(-6, 7)
    cells = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 1, 3)
    pts = ((cells + _diamond(a)[None]) * a).reshape(-1, 3) @ np.stack([x, y, z]).T
    eps = 1e-6
    sure = pts[((pts > eps) & (pts < size - eps)).all(1)]
    maybe = pts[((pts > -eps) & (pts < size + eps)).all(1)]
    from scipy.spatial import cKDTree
    assert (cKDTree(got).query(sure)[0] < 1e-4).all()         # every interior lattice point is present
    assert (cKDTree(maybe).query(got)[0] < 1e-4).all()        # and nothing else
    assert len(sure) <= len(got) <= len(maybe)
    assert (sp.supercell_sites["atomic_number"] == 14).all() and (sp.supercell_sites["DW"] == np.float32(0.07)).all()


# --- pieces -------------------------------------------------------------------------------------------------------


def test_lattice_matrix_convention():
    p = (5.62367184, 8.81191945, 11.49202588, 87.19746410, 75.95579729, 78.51700298)
    m = S.lattice_matrix(*p)
    np.testing.assert_allclose(m, O.lattice_from_parameters(*p), atol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), p[:3], rtol=1e-14)
    np.testing.assert_allclose(S._angles_deg(m.T), p[3:], rtol=1e-12)
    assert m[2, 0] == 0 and m[2, 1] == 0 and m[0, 1] == 0            # c along z, a in the xz plane
    assert np.allclose(S.lattice_matrix(3, 4, 5, 90, 90, 90), np.diag([3, 4, 5]), atol=1e-15)


def test_symmetry_operations():
    si_ops = S.read_cif(SI)["_space_group_symop_operation_xyz"]
    for s in si_ops + ["1/2+x, -y+z, x-y+1/3", "X, Y, Z", "2x, x/2, -z+0.25", "+z, -x, y-1/4"]:
        np.testing.assert_allclose(S.symop_from_xyz(s), O.symop(s), atol=1e-15, err_msg=s)
    w = S.symop_from_xyz("y, -x, z")                                   # 4-fold rotation, twice = 2-fold
    np.testing.assert_allclose(w @ w, S.symop_from_xyz("-x, -y, z"))
    t = S.symop_from_xyz("x+1/2, y, z") @ S.symop_from_xyz("-x, y+1/4, z")
    np.testing.assert_allclose(t, S.symop_from_xyz("-x+1/2, y+1/4, z"))
    with pytest.raises(ValueError):
        S.symop_from_xyz("x, y")


def test_cif_reader_forms(tmp_path):
    text = """# comment line
data_first
_cell_length_a 4.0(2)
_cell_length_b   4.0
_cell_length_c 6.5 # trailing comment
_cell_angle_alpha 90
_cell_angle_beta 90
_cell_angle_gamma 120
_symmetry_space_group_name_H-M 'P 1'
_note
;
multi-line text field
with 'quotes'
;
loop_
_atom_site_label
_atom_site_type_symbol
_atom_site_fract_x
_atom_site_fract_y
_atom_site_fract_z
_atom_site_occupancy
Ga1 Ga3+ 0.3333 0.6667 0.0 1.0
N1 'N' 0.3333 0.6667 0.377(3) 0.5
data_second
_cell_length_a 99
"""
    p = tmp_path / "gan.cif"
    p.write_text(text)
    d = S.read_cif(str(p))
    assert d["_cell_length_a"] == "4.0(2)" and d["_note"].startswith("multi-line")
    assert d["_atom_site_type_symbol"] == ["Ga3+", "N"]
    sp = S.SupercellBuilder(str(p))
    assert len(sp.sym_ops) == 1 and np.allclose(sp.sym_ops[0], np.eye(4))      # P1 without an operation list
    assert list(sp.sites["atomic_number"]) == [31, 7]
    np.testing.assert_allclose(sp.sites["frac_z"], [0.0, 0.377], rtol=1e-6)
    np.testing.assert_allclose(sp.sites["occ"], [1.0, 0.5])
    np.testing.assert_allclose(sp.lattice_const, [4.0, 4.0, 6.5])
    np.testing.assert_allclose(sp.lattice_angles, [90, 90, 120], atol=1e-12)


def test_errors_and_no_mutation():
    with pytest.raises(NotImplementedError):
        S.SupercellBuilder(MVC, primitive=True)
    with pytest.raises(ValueError):
        S.atomic_number("Xx")
    assert S.atomic_number("Fe") == 26 and S.atomic_number("O2-") == 8 and S.atomic_number("Lr") == 103
    sp = S.SupercellBuilder(MVC)
    scaling = np.array([1, 1, 1])
    sp.transform_unit_cell(centering_type="C", scaling=scaling)
    assert list(scaling) == [1, 1, 1]
    sp.build_unit_cell()
    with pytest.raises(ValueError):
        sp.make_orthogonal_supercell(projection_rule="other")


@pytest.mark.parametrize("cen,det", [("I", 4), ("C", 2), ("R", 9), ("F", 2)])
def test_centerings_scale_the_site_count(cen, det):
    """The site search finds |det P| times the cell's 45 sites, as the
    oracle's restatement does (positions to the 6-decimal rounding)."""
    sp = S.SupercellBuilder(MVC)
    b0, ops0, s0 = sp.basis.copy(), list(sp.sym_ops), sp.sites.copy()
    sp.transform_unit_cell(centering_type=cen)
    assert sp.sites.size == 45 * det
    _, lc, ang, _, sites = O.transform(b0, ops0, s0, centering=cen)
    np.testing.assert_allclose(sp.lattice_const, lc, rtol=1e-14)
    np.testing.assert_allclose(sp.lattice_angles, ang, rtol=1e-12)
    assert (sp.sites["atom_type"] == sites["atom_type"]).all()
    for f in ("frac_x", "frac_y", "frac_z"):
        assert np.abs(sp.sites[f] - sites[f]).max() < 2e-6


def test_xyz_files_round_trip(tmp_path):
    """to_XYZ writes the Kirkland-style file MSA reads (header '#', the
    extents line with the x maximum twice, records, '-1'); reading it the way
    MSA does returns the records."""
    sp = S.SupercellBuilder(SI)
    sp.build_unit_cell()
    sp.make_orthogonal_supercell(supercell_size=np.array([12.0, 8.0, 6.0]))
    path = tmp_path / "si.XYZ"
    assert sp.to_XYZ(str(path), supercell=True)
    lines = path.read_text().splitlines()
    x = sp.supercell_sites["x"]
    assert lines[0] == "#" and lines[-1] == "-1"
    assert lines[1] == "      %2.4f %2.4f %2.4f" % (x.max(), x.max(), sp.supercell_sites["z"].max())
    back = np.genfromtxt(str(path), skip_footer=1, skip_header=2, dtype=S.XYZ_FILE_DTYPE)
    assert (back["atomic_number"] == 14).all() and len(back) == len(sp.supercell_sites)
    assert np.abs(back["x"] - x).max() <= 5e-5 and np.abs(back["DW"] - 0.07).max() < 1e-6
    from namsa_amd.msa import MSAGPU
    arr = MSAGPU._load_supercell(str(path))
    assert len(arr) == len(back)
    assert sp.to_xyz(str(tmp_path / "uc.xyz"))                      # unit cell, Cartesian
    assert (tmp_path / "uc.xyz").read_text().splitlines()[0] == "8"


GAN_CIF = """data_GaN
_cell_length_a 3.189
_cell_length_b 3.189
_cell_length_c 5.185
_cell_angle_alpha 90
_cell_angle_beta 90
_cell_angle_gamma 120
_symmetry_space_group_name_H-M 'P 63 m c'
loop_
_symmetry_equiv_pos_as_xyz
'x, y, z'
'-y, x-y, z'
'-x+y, -x, z'
'-x, -y, z+1/2'
'y, -x+y, z+1/2'
'x-y, x, z+1/2'
'-y, -x, z'
'-x+y, y, z'
'x, x-y, z'
'y, x, z+1/2'
'x-y, -y, z+1/2'
'-x, -x+y, z+1/2'
loop_
_atom_site_label
_atom_site_type_symbol
_atom_site_fract_x
_atom_site_fract_y
_atom_site_fract_z
_atom_site_occupancy
Ga1 Ga 0.333333 0.666667 0.0 1
N1 N 0.333333 0.666667 0.377 1
"""


def _wurtzite_frac(u=0.377):
    return np.array([[1 / 3, 2 / 3, 0.0], [2 / 3, 1 / 3, 0.5], [1 / 3, 2 / 3, u], [2 / 3, 1 / 3, 0.5 + u]])


def test_hexagonal_unit_cell_is_wurtzite(tmp_path):
    """P6_3mc operations on the two GaN sites give the four wurtzite
    positions (Ga at 2b, N at 2b + u)."""
    p = tmp_path / "gan.cif"
    p.write_text(GAN_CIF)
    sp = S.SupercellBuilder(str(p))
    sp.build_unit_cell()
    uc = sp.unit_cell_positions
    frac = np.column_stack([uc["frac_x"], uc["frac_y"], uc["frac_z"]]).astype(np.float64)
    want = _wurtzite_frac()
    assert list(uc["atom_type"]) == ["Ga", "Ga", "N", "N"]
    got = sorted(map(tuple, np.round(frac, 5)))
    assert got == sorted(map(tuple, np.round(want, 5)))
    np.testing.assert_allclose(sp.lattice_angles, [90, 90, 120], atol=1e-12)


def test_hexagonal_supercell_against_lattice_enumeration(tmp_path):
    """A non-orthogonal cell: the supercell equals the wurtzite crystal
    (lattice points + basis, Cartesian), rotated into the box frame and
    enumerated directly, for the lattice-vector projection rule. Also
    checks that the box frame is a rotation and the nearest Ga-N distance."""
    p = tmp_path / "gan.cif"
    p.write_text(GAN_CIF)
    sp = S.SupercellBuilder(str(p))
    sp.build_unit_cell()
    size = np.array([12.0, 10.0, 11.0])
    sp.make_orthogonal_supercell(supercell_size=size, projec_1=np.array([1, 0, 0]), projec_2=np.array([0, 0, 1]),
                                 projection_rule="lattice")
    got = _pos(sp.supercell_sites)
    T, _, _, _, _, _ = sp._supercell_frame(size, [1, 0, 0], [0, 0, 1], "lattice")
    cart = sp.CartesianBasis()
    R = T @ np.linalg.inv(cart / np.linalg.norm(cart, axis=0))      # box frame from Cartesian
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    n = np.arange(-8, 9)
    cells = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    pts = ((cells[:, None, :] + _wurtzite_frac()[None]) @ cart.T).reshape(-1, 3) @ R.T
    eps = 1e-6
    sure = pts[((pts > eps) & (pts < size - eps)).all(1)]
    maybe = pts[((pts > -eps) & (pts < size + eps)).all(1)]
    from scipy.spatial import cKDTree
    assert (cKDTree(got).query(sure)[0] < 1e-4).all()
    assert (cKDTree(maybe).query(got)[0] < 1e-4).all()
    ga, nn = got[sp.supercell_sites["atom_type"] == "Ga"], got[sp.supercell_sites["atom_type"] == "N"]
    a, c, u = 3.189, 5.185, 0.377                  # wurtzite bonds: u c along c, sqrt(a^2/3 + (1/2 - u)^2 c^2) basal
    assert abs(cKDTree(nn).query(ga)[0].min() - min(u * c, np.sqrt(a * a / 3 + ((0.5 - u) * c) ** 2))) < 1e-4


def test_primitive_fcc_transform_closed_form():
    """transform_unit_cell('F', primitive=True) on diamond Si: the rhombohedral
    primitive cell, a/sqrt(2) edges at 60 degrees, volume a^3/4, two atoms at 0
    and a/4 (1, 1, 1); the oracle's restatement agrees."""
    a = 5.43094
    sp = S.SupercellBuilder(SI)
    b0, ops0, s0 = sp.basis.copy(), list(sp.sym_ops), sp.sites.copy()
    sp.transform_unit_cell(centering_type="F", primitive=True)
    sp.build_unit_cell()
    np.testing.assert_allclose(sp.lattice_const, [a / np.sqrt(2)] * 3, rtol=1e-12)
    np.testing.assert_allclose(sp.lattice_angles, [60.0] * 3, rtol=1e-12)
    assert abs(abs(sp.lattice_volume) - a ** 3 / 4) < 1e-9
    np.testing.assert_allclose(np.sort(_pos(sp.xyz_positions), axis=0), [[0, 0, 0], [a / 4] * 3], atol=1e-5)
    basis, _, _, ops, sites = O.transform(b0, ops0, s0, centering="F", primitive=True)
    _, cart, _ = O.unit_cell(basis, ops, sites)
    np.testing.assert_allclose(cart, _pos(sp.xyz_positions), atol=1e-5)
