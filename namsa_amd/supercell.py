"""Structure front-end: CIF file -> symmetry-expanded unit cell -> oriented
orthogonal supercell, giving the site records that MSAGPU takes
(namsa/supercell.py).

The reference reads the CIF through pymatgen (absent here). This module
parses the file itself and builds the lattice matrix as pymatgen's
`Lattice.from_parameters` does, so the basis vectors are the same. It reads
the symmetry operations from the CIF's xyz strings, as pymatgen's
`SymmOp.from_xyz_str` does. After that, every step follows
namsa/supercell.py. The per-site Python loops are vectorised: the
translations of all sites are tiled at once, and the final ordering is one
lexsort, so a supercell of a million atoms takes a fraction of a second.

Host-side code. The records are consumed by `MSAGPU(..., supercell=...)`,
whose slice builder runs on the GPU.
"""
import os
import re
from warnings import warn

import numpy as np

# numpy structured dtypes of namsa/supercell.py:69-74
SITES_DTYPE = [("atom_type", "|U16"), ("atomic_number", "i"), ("frac_x", "f"), ("frac_y", "f"),
               ("frac_z", "f"), ("occ", "f"), ("DW", "f")]
SUPERCELL_SITES_DTYPE = [("atom_type", "|U16"), ("atomic_number", "i"), ("x", "f"), ("y", "f"), ("z", "f"),
                         ("occ", "f"), ("DW", "f")]
XYZ_DTYPE = [("atom_type", "|U16"), ("x", "f"), ("y", "f"), ("z", "f")]
XYZ_FILE_DTYPE = [("atomic_number", "i"), ("x", "f"), ("y", "f"), ("z", "f"), ("occ", "f"), ("DW", "f")]

#: Debye-Waller value every site receives (namsa/supercell.py:103-109: the
#: B/U columns are never used, so the constant always applies).
DW_DEFAULT = 0.07

ELEMENTS = (
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K",
    "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb",
    "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs",
    "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta",
    "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa",
    "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt",
    "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og")
_Z_OF = {s: i + 1 for i, s in enumerate(ELEMENTS)}


def atomic_number(symbol):
    """Z of an element symbol ('Fe'). An oxidation-state suffix ('Fe3+',
    'O2-') is accepted and dropped; pymatgen's Element() refuses it."""
    m = re.match(r"([A-Z][a-z]?)", symbol.strip())
    if not m or m.group(1) not in _Z_OF:
        raise ValueError(f"unknown element symbol {symbol!r}")
    return _Z_OF[m.group(1)]


# ----------------------------------------------------------------------------------------------------------------
# CIF reading

_TOKEN = re.compile(r"""'(.*?)'(?=\s|$)|"(.*?)"(?=\s|$)|(\S+)""")


def _tokens(text):
    """Yield the tokens of a CIF: quoted and bare values, tags, loop_ and
    data_ keywords, and semicolon text fields as one value each."""
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i]
        if line.startswith(";"):
            buf = [line[1:]]
            i += 1
            while i < len(lines) and not lines[i].startswith(";"):
                buf.append(lines[i])
                i += 1
            i += 1
            yield "\n".join(buf).strip(), True
            continue
        for m in _TOKEN.finditer(line):
            if m.group(3) is not None and m.group(3).startswith("#"):
                break
            quoted = m.group(3) is None
            yield (m.group(1) if m.group(1) is not None else m.group(2) if m.group(2) is not None
                   else m.group(3)), quoted
        i += 1


def read_cif(path_or_text):
    """The first data block of a CIF as {tag: value or list of values} (all
    strings, tags lower-cased), as the first entry of pymatgen's
    CifParser.as_dict() (namsa/supercell.py:95-96)."""
    text = os.fspath(path_or_text)
    if "\n" not in text:
        with open(text) as f:
            text = f.read()
    data, seen_block = {}, False
    toks = list(_tokens(text))
    i = 0
    while i < len(toks):
        tok, quoted = toks[i]
        low = tok.lower()
        if not quoted and low.startswith("data_"):
            if seen_block:
                break
            seen_block = True
            i += 1
        elif not quoted and low == "loop_":
            i += 1
            tags = []
            while i < len(toks) and not toks[i][1] and toks[i][0].startswith("_"):
                tags.append(toks[i][0].lower())
                i += 1
            vals = []
            while i < len(toks):
                t, q = toks[i]
                tl = t.lower()
                if not q and (t.startswith("_") or tl == "loop_" or tl.startswith("data_")
                              or tl.startswith("save_") or tl == "global_"):
                    break
                vals.append(t)
                i += 1
            if not tags or len(vals) % len(tags):
                raise ValueError(f"malformed CIF loop over {tags}: {len(vals)} values")
            for k, tag in enumerate(tags):
                data[tag] = vals[k::len(tags)]
        elif not quoted and tok.startswith("_"):
            if i + 1 >= len(toks):
                raise ValueError(f"CIF tag {tok} has no value")
            data[low] = toks[i + 1][0]
            i += 2
        else:
            i += 1
    if not seen_block:
        raise ValueError("no data_ block in CIF")
    return data


def cif_float(s):
    """A CIF number, with a standard uncertainty '(n)' dropped."""
    return float(re.sub(r"\(\d+\)$", "", s.strip()))


_TERM = re.compile(r"([+-]?)([^+-]+)")


def _number(s):
    if "/" in s:
        n, d = s.split("/")
        return float(n) / float(d)
    return float(s)


def symop_from_xyz(xyz):
    """4x4 affine matrix of a symmetry operation written as 'x+1/2, -y, z'
    (the form pymatgen's SymmOp.from_xyz_str reads)."""
    parts = xyz.lower().replace(" ", "").split(",")
    if len(parts) != 3:
        raise ValueError(f"symmetry operation {xyz!r} does not have three components")
    m = np.eye(4)
    m[:3, :3] = 0.0
    for row, part in enumerate(parts):
        if not part:
            raise ValueError(f"empty component in symmetry operation {xyz!r}")
        for sign, body in _TERM.findall(part):
            sgn = -1.0 if sign == "-" else 1.0
            var = re.search(r"[xyz]", body)
            if var:
                pre, post = body[:var.start()].rstrip("*"), body[var.end():]
                coef = _number(pre) if pre else 1.0
                if post.startswith("/"):
                    coef /= float(post[1:])
                elif post.startswith("*"):
                    coef *= _number(post[1:])
                elif post:
                    raise ValueError(f"cannot read term {body!r} of symmetry operation {xyz!r}")
                m[row, "xyz".index(var.group(0))] += sgn * coef
            else:
                m[row, 3] += sgn * _number(body)
    return m


_SYMOP_TAGS = ("_symmetry_equiv_pos_as_xyz", "_symmetry_equiv_pos_as_xyz_",
               "_space_group_symop_operation_xyz", "_space_group_symop_operation_xyz_")


def cif_symops(data):
    """Affine symmetry operations of a CIF block, from the first of the xyz
    lists pymatgen looks for. A P1 file without a list gets the identity;
    other space groups need the list (no space-group tables are bundled)."""
    for tag in _SYMOP_TAGS:
        if tag in data:
            xyz = data[tag]
            return [symop_from_xyz(s) for s in ([xyz] if isinstance(xyz, str) else xyz)]
    hm = str(data.get("_symmetry_space_group_name_h-m", data.get("_space_group_name_h-m_alt", ""))).replace(" ", "")
    num = str(data.get("_symmetry_int_tables_number", data.get("_space_group_it_number", "")))
    if hm == "P1" or num == "1":
        return [np.eye(4)]
    raise NotImplementedError("the CIF lists no symmetry operations; space-group tables are not bundled, so "
                              "add a _symmetry_equiv_pos_as_xyz loop")


def lattice_matrix(a, b, c, alpha, beta, gamma):
    """Rows = lattice vectors, in the convention of pymatgen's
    Lattice.from_parameters (c along z, a in the xz plane)."""
    al, be, ga = np.radians([alpha, beta, gamma])
    cos_al, cos_be, cos_ga = np.cos([al, be, ga])
    sin_al, sin_be = np.sin(al), np.sin(be)
    val = np.clip((cos_al * cos_be - cos_ga) / (sin_al * sin_be), -1.0, 1.0)
    gamma_star = np.arccos(val)
    return np.array([[a * sin_be, 0.0, a * cos_be],
                     [-b * sin_al * np.cos(gamma_star), b * sin_al * np.sin(gamma_star), b * cos_al],
                     [0.0, 0.0, float(c)]])


def _angles_deg(basis):
    """(alpha, beta, gamma) in degrees of the column vectors of `basis`."""
    v = basis.T
    n = np.linalg.norm(v, axis=1)
    return np.degrees(np.array([np.arccos(np.dot(v[j], v[k]) / (n[j] * n[k])) for j, k in ((1, 2), (2, 0), (0, 1))]))


# ----------------------------------------------------------------------------------------------------------------
# Builder


class Centerings:
    """Cell transformations for a lattice centering (namsa/supercell.py:8-53):
    conventional matrices, or half (a third for R) of them with primitive=True."""

    _CONVENTIONAL = {
        "I": [[-1, 1, 1], [1, -1, 1], [1, 1, -1]],
        "A": [[1, 0, 0], [0, 1, -1], [0, 1, 1]],
        "C": [[1, 1, 0], [-1, 1, 0], [0, 0, 1]],
        "R": [[2, -1, -1], [1, 1, -2], [1, 1, 1]],
        "F": [[0, 1, 1], [1, 0, 1], [1, 1, 0]],
    }

    def __init__(self, centering="P", primitive=False):
        self.centering = centering
        self.transformations = {"P": np.array(np.identity(3))}
        for k, m in self._CONVENTIONAL.items():
            m = np.array(m)
            self.transformations[k] = (m / 3.0 if k == "R" else 0.5 * m) if primitive else m

    @property
    def trans_mat(self):
        return self.transformations[self.centering]


class SupercellBuilder:
    """Build a supercell from a CIF file (namsa/supercell.py:56-464).

    Attributes and methods keep the reference's names and meaning: basis
    (lattice vectors as columns), lattice_const, lattice_angles (degrees),
    sym_ops (4x4 affine), sites, unit_cell_positions, xyz_positions,
    supercell_sites (sorted records for MSAGPU) and supercell_xyz_positions.
    """

    def __init__(self, cif_file, primitive=False, debug=False, verbose=False):
        """Parse the CIF (namsa/supercell.py:60-117).

        primitive=True asks pymatgen for the primitive standard cell, which
        needs a symmetry finder; it is not supported and raises. Every site
        gets DW = 0.07, as in the reference, whose B/U-column read always
        falls through to that constant.
        """
        if primitive:
            raise NotImplementedError("primitive=True needs a space-group finder (pymatgen/spglib); reduce the "
                                      "cell before writing the CIF, or use transform_unit_cell(primitive=True)")
        self.debug, self.verbose = debug, verbose
        self.sites_dtype, self.supercell_sites_dtype = SITES_DTYPE, SUPERCELL_SITES_DTYPE
        self.xyz_dtype, self.XYZ_dtype = XYZ_DTYPE, XYZ_FILE_DTYPE
        self.p_mat, self.P_mat, self.Q_mat = np.identity(3), np.identity(4), np.identity(4)

        data = read_cif(str(cif_file))
        self.cif_data = data
        abc = [cif_float(data["_cell_length_" + k]) for k in "abc"]
        ang = [cif_float(data["_cell_angle_" + k]) for k in ("alpha", "beta", "gamma")]
        self.basis = lattice_matrix(*abc, *ang).T
        self.lattice_const = np.linalg.norm(self.basis, axis=0)
        self.lattice_angles = _angles_deg(self.basis)
        self.lattice_volume = abs(np.linalg.det(self.basis))
        self.sym_ops = cif_symops(data)

        def column(tag):
            v = data[tag]
            return [v] if isinstance(v, str) else list(v)

        frac = np.array([[cif_float(s) for s in column("_atom_site_fract_" + k)] for k in "xyz"]).T
        if "_atom_site_type_symbol" in data:
            symbols = column("_atom_site_type_symbol")
        else:       # the reference needs the type column; fall back to the labels' element part
            symbols = [re.match(r"[A-Za-z]+", s).group(0) for s in column("_atom_site_label")]
        occ = ([cif_float(s) for s in column("_atom_site_occupancy")] if "_atom_site_occupancy" in data
               else [1.0] * len(symbols))
        sites = np.zeros(len(symbols), dtype=SITES_DTYPE)
        sites["atom_type"] = symbols
        sites["atomic_number"] = [atomic_number(s) for s in symbols]
        sites["frac_x"], sites["frac_y"], sites["frac_z"] = frac.T
        sites["occ"] = occ
        sites["DW"] = DW_DEFAULT
        self.sites = sites
        self.print_verbose(" Lattice vectors:\n", np.round(self.basis, 4))
        self.print_verbose(" Lattice Constants (Å): \n", np.round(self.lattice_const, 4))
        self.print_verbose(" Lattice Angles (deg.): \n", np.round(self.lattice_angles, 4))
        self.print_verbose(" Volume (Å**3): \n", np.round(self.lattice_volume, 4))

    def print_debug(self, *args, **kwargs):
        if self.debug:
            print(*args, **kwargs)

    def print_verbose(self, *args, **kwargs):
        if self.verbose:
            print(*args, **kwargs)

    # -- cell transformation ------------------------------------------------------------------------------------

    def transform_unit_cell(self, centering_type="P", primitive=False, scaling=np.array([1, 1, 1]), rotation=None,
                            origin_shift=np.array([0, 0, 0]), max_iter=4):
        """Change the cell by P = (rotation or the centering matrix) with its
        columns scaled by `scaling` and the origin moved by `origin_shift`
        (namsa/supercell.py:158-244). The basis becomes basis·P, each
        symmetry operation becomes Q·W·P with Q = P⁻¹, and, for a centering
        other than P, the sites of the new cell are searched for. The search
        tiles every site over a growing range of translations until the count
        stops changing or max_iter is reached. It keeps positions that are
        distinct to 2 decimals, rounded to 6 decimals and below 1.

        Unlike the reference, the caller's `scaling` array is not modified.
        """
        scaling = np.array(scaling)
        p = np.array(rotation if rotation is not None else Centerings(centering_type, primitive=primitive).trans_mat)
        for k in range(3):
            p[:, k] = p[:, k] * scaling[k]
        self.p_mat = p
        self.P_mat = np.identity(4)
        self.P_mat[:3, :3] = p
        self.P_mat[:3, 3] = origin_shift
        self.Q_mat = np.linalg.inv(self.P_mat)

        self.basis = np.dot(self.basis, p)
        self.lattice_const = np.linalg.norm(self.basis, axis=0)
        self.lattice_angles = _angles_deg(self.basis)
        # the reference's triple product of the basis rows (namsa/supercell.py:191)
        self.lattice_volume = np.dot(self.basis[0], np.cross(self.basis[1], self.basis[2]))
        self.print_verbose("Transformed Lattice vectors:\n", np.round(self.basis, 5))
        self.sym_ops = [np.dot(self.Q_mat, w).dot(self.P_mat) for w in self.sym_ops]

        if centering_type == "P":
            return
        n = scaling.astype(int)
        previous, it = 0, 0
        while True:
            found = self._search_sites(self.sites, self.Q_mat, n)
            self.print_debug("Iteration #%d: found %d atomic sites, previous %d" % (it, found.size, previous))
            if found.size == previous or it >= max_iter:
                self.sites = found
                break
            previous, it = found.size, it + 1
            n = n + 1
        self.print_verbose("Total number of atomic sites: %d" % self.sites.size)

    @staticmethod
    def _search_sites(sites, Q, n):
        """Sites of the cell Q maps onto: each site tiled over translations
        0..n per axis, mapped by Q, wrapped into [0, 1), deduplicated at 2
        decimals (first occurrence kept), rounded to 6 decimals, and kept
        when every coordinate is below 1 (namsa/supercell.py:211-230)."""
        n_x, n_y, n_z = (int(v) for v in n)
        out = []
        for s in sites:
            gx = np.linspace(s["frac_x"], s["frac_x"] + n_x, n_x + 1)
            gy = np.linspace(s["frac_y"], s["frac_y"] + n_y, n_y + 1)
            gz = np.linspace(s["frac_z"], s["frac_z"] + n_z, n_z + 1)
            X, Y, Z = (g.ravel() for g in np.meshgrid(gx, gy, gz))
            pos = np.dot(Q, np.stack([X, Y, Z, np.ones_like(Z)]))[:3].T
            pos -= pos // 1
            _, first = np.unique(np.round(pos, 2), axis=0, return_index=True)
            pos = np.round(pos[first], 6)
            pos = pos[(pos < 1.0).all(axis=1)]
            rec = np.zeros(len(pos), dtype=SITES_DTYPE)
            for f in ("atom_type", "atomic_number", "occ", "DW"):
                rec[f] = s[f]
            rec["frac_x"], rec["frac_y"], rec["frac_z"] = pos.T
            out.append(rec)
        return np.concatenate(out) if out else np.zeros(0, dtype=SITES_DTYPE)

    # -- unit cell ----------------------------------------------------------------------------------------------

    def build_unit_cell(self, xyz=None):
        """Apply every symmetry operation to every site, wrap into [0, 1) and
        keep the positions distinct to 2 decimals (first occurrence, in the
        sorted order of the rounded coordinates), per site in file order
        (namsa/supercell.py:246-283). Sets unit_cell_positions (fractional)
        and xyz_positions (Cartesian, basis · frac)."""
        ops = np.array(self.sym_ops)[:, :3, :]                       # (n_ops, 3, 4)
        frac_l, rows = [], []
        for k, s in enumerate(self.sites):
            site = np.array([s["frac_x"], s["frac_y"], s["frac_z"], 1.0])
            pos = ops @ site
            pos -= pos // 1
            _, first = np.unique(np.round(pos, 2), return_index=True, axis=0)
            frac_l.append(pos[first])
            rows.append(np.full(len(first), k))
        frac = np.concatenate(frac_l) if frac_l else np.zeros((0, 3))
        src = self.sites[np.concatenate(rows)] if rows else self.sites[:0]
        cell = np.zeros(len(frac), dtype=SITES_DTYPE)
        for f in ("atom_type", "atomic_number", "occ", "DW"):
            cell[f] = src[f]
        cell["frac_x"], cell["frac_y"], cell["frac_z"] = frac.T
        self.unit_cell_positions = cell
        cart = np.dot(self.basis, frac.T).T
        xyz_arr = np.zeros(len(frac), dtype=XYZ_DTYPE)
        xyz_arr["atom_type"] = src["atom_type"]
        xyz_arr["x"], xyz_arr["y"], xyz_arr["z"] = cart.T
        self.xyz_positions = xyz_arr
        if xyz is not None:
            self.to_xyz(xyz, supercell=False)

    # -- orientation and supercell ------------------------------------------------------------------------------

    def CartesianBasis(self):
        """Upper-triangular basis from the lattice constants and angles
        (namsa/supercell.py:321-335): a along x, b in the xy plane."""
        a, b, c = self.lattice_const
        al, be, ga = np.radians(self.lattice_angles)
        c_x = c * np.cos(be)
        c_y = c * np.cos(al) * np.sin(ga)
        return np.array([[a, b * np.cos(ga), c_x],
                         [0, b * np.sin(ga), c_y],
                         [0, 0, np.sqrt(np.abs(c ** 2 - c_x ** 2 - c_y ** 2))]])

    def OrientationMatrix(self, hkl=np.array([0, 0, 1]), uvw=np.array([0, 1, 0]), mode="uvw"):
        """Set R, whose rows are the new x, y, z axes (namsa/supercell.py:
        337-366). mode 'uvw' takes z along the direct-lattice vector uvw, mode
        'hkl' takes z along the reciprocal vector G_hkl, and 'none' gives the
        identity."""
        a1, a2, a3 = self.basis.T
        UVW = uvw[0] * a1 + uvw[1] * a2 + uvw[2] * a3
        recip = np.array([np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)]) / self.lattice_volume
        HKL = hkl[0] * recip[0] + hkl[1] * recip[1] + hkl[2] * recip[2]
        if mode == "none":
            self.R = np.identity(3)
            return
        if mode == "hkl":
            x_ax, z_ax = np.cross(UVW, HKL), HKL
        elif mode == "uvw":
            x_ax, z_ax = np.cross(HKL, UVW), UVW
        else:
            raise ValueError(f"mode must be 'uvw', 'hkl' or 'none', not {mode!r}")
        x_ax = x_ax / np.linalg.norm(x_ax)
        z_ax = z_ax / np.linalg.norm(z_ax)
        y_ax = np.cross(z_ax, x_ax)
        y_ax = y_ax / np.linalg.norm(y_ax)
        self.R = np.vstack([x_ax, y_ax, z_ax])
        self.print_verbose("Orientation Matrix:\n", np.round(self.R, 4))

    def _supercell_frame(self, supercell_size, projec_1, projec_2, projection_rule):
        """trans_mat, the translation vectors and the range of cell repeats of
        make_orthogonal_supercell (namsa/supercell.py:368-429)."""
        cart = self.CartesianBasis()
        cart_unit = cart / np.linalg.norm(cart, axis=0)
        if projection_rule == "components":
            # each projection vector weights the components of the summed basis columns (element-wise), as
            # namsa/supercell.py:377-380 computes it
            row_sum = cart.sum(axis=1)
            b_new, c_new = np.asarray(projec_1) * row_sum, np.asarray(projec_2) * row_sum
        elif projection_rule == "lattice":
            # the direct-lattice vectors sum_i p_i a_i
            b_new, c_new = cart @ np.asarray(projec_1, float), cart @ np.asarray(projec_2, float)
        else:
            raise ValueError(f"projection_rule must be 'components' or 'lattice', not {projection_rule!r}")
        b_new = b_new / np.linalg.norm(b_new)
        c_new = c_new / np.linalg.norm(c_new)
        a_new = np.cross(b_new, c_new)
        if np.round(np.dot(b_new, c_new), 5) != 0.0:
            b_new = np.cross(a_new, c_new)
            if np.round(np.dot(b_new, c_new), 5) != 0.0:
                self.print_debug("Can not find orthonormal basis")
        new_basis = np.column_stack([a_new, b_new, c_new])
        trans_mat = np.dot(cart_unit.T, new_basis).T
        size = np.asarray(supercell_size, dtype=float)
        corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=float) * size
        corners_org = np.dot(np.linalg.inv(trans_mat), corners.T).T
        hi = np.ceil(corners_org.max(0) / self.lattice_const).astype(int)
        lo = np.floor(corners_org.min(0) / self.lattice_const).astype(int)
        t_vec = np.dot(trans_mat, np.diag(self.lattice_const))         # columns: the translation vectors
        return trans_mat, t_vec, lo, hi, np.linalg.norm(cart, axis=0), size

    def make_orthogonal_supercell(self, supercell_size=np.array([20.0, 20.0, 20.0]), projec_1=np.array([1, 0, 0]),
                                  projec_2=np.array([0, 0, 1]), xyz=None, XYZ=None, projection_rule="components"):
        """Fill an orthogonal box of `supercell_size` Å (namsa/supercell.py:
        368-464) with the unit cell's sites and set supercell_sites, sorted by
        (atom_type, atomic_number, x, y, z, occ, DW).

        The new y and z axes follow projec_1 and projec_2. With
        projection_rule="components" (the reference's code), an axis is the
        projection vector times the sum of the Cartesian basis columns, taken
        element by element. With "lattice", an axis is the lattice vector
        sum_i p_i a_i; the reference's own fixture
        cif_files/supercell_mvc-9418.xyz was made by this rule. The two rules
        agree whenever the Cartesian basis is diagonal (cubic, tetragonal and
        orthorhombic cells).

        As in the reference, each unit-cell site is placed at trans_mat ·
        (frac ⊙ |cell axes|). trans_mat holds the normalised cell axes, so
        this is the site's Cartesian position rotated into the box frame. It is
        translated by trans_mat · diag(a, b, c) (the lattice vectors in that
        frame) times every integer triple in the repeat range, and kept when
        0 ≤ x, y, z ≤ size (bounds inclusive, in float64, before the float32
        records). The repeat range spans the fractional coordinates of the box
        corners, floored and ceiled, which covers every lattice point of the
        box for any cell shape.
        """
        trans_mat, t_vec, lo, hi, bnorm, size = self._supercell_frame(supercell_size, projec_1, projec_2,
                                                                      projection_rule)
        self.print_debug("Supercell %s contains unit cell repetition (a,b,c): %s" % (size, hi - lo))
        uc = self.unit_cell_positions
        frac = np.column_stack([uc["frac_x"], uc["frac_y"], uc["frac_z"]]).astype(np.float64)
        base = (trans_mat @ (frac * bnorm).T).T                          # (n_sites, 3)
        cx, cy, cz = (np.arange(lo[k], hi[k] + 1, dtype=float) for k in range(3))
        cells = np.stack([g.ravel() for g in np.meshgrid(cx, cy, cz, indexing="ij")], axis=1)
        shifts = cells @ t_vec.T                                         # (n_cells, 3)
        keep_site, keep_pos = [], []
        chunk = max(1, (1 << 24) // max(1, len(shifts)))               # bounds the (sites x cells x 3) temporary
        for s0 in range(0, len(base), chunk):
            pos = base[s0:s0 + chunk, None, :] + shifts[None, :, :]
            inside = ((pos >= 0) & (pos <= size)).all(axis=2)
            si, ci = np.nonzero(inside)
            keep_site.append(si + s0)
            keep_pos.append(pos[si, ci])
        site_idx = np.concatenate(keep_site) if keep_site else np.zeros(0, int)
        pos = np.concatenate(keep_pos) if keep_pos else np.zeros((0, 3))
        rec = np.zeros(len(pos), dtype=SUPERCELL_SITES_DTYPE)
        for f in ("atom_type", "atomic_number", "occ", "DW"):
            rec[f] = uc[f][site_idx]
        rec["x"], rec["y"], rec["z"] = pos.T
        order = np.lexsort((rec["DW"], rec["occ"], rec["z"], rec["y"], rec["x"], rec["atomic_number"],
                            rec["atom_type"]))
        self.supercell_sites = rec[order]
        sc = np.zeros(len(rec), dtype=XYZ_DTYPE)
        for f in ("atom_type", "x", "y", "z"):
            sc[f] = self.supercell_sites[f]
        self.supercell_xyz_positions = sc
        if xyz is not None:
            self.to_xyz(xyz, supercell=True)
            self.print_verbose("saving xyz file %s" % xyz)
        if XYZ is not None:
            self.to_XYZ(XYZ, supercell=True)
            self.print_verbose("saving XYZ file %s" % XYZ)

    # -- files --------------------------------------------------------------------------------------------------

    def to_xyz(self, filepath, supercell=False):
        """Write an xyz file: the atom count, a blank line, then
        'symbol    x    y    z' with 4 decimals (namsa/supercell.py:286-298).
        supercell=False writes the Cartesian unit cell (the reference raises
        there, reading x from the fractional records)."""
        cell = self.supercell_xyz_positions if supercell else self.xyz_positions
        lines = [f"{len(cell)}", ""]
        lines += ["%s    %2.4f    %2.4f    %2.4f" % (t, x, y, z)
                  for t, x, y, z in zip(cell["atom_type"], cell["x"], cell["y"], cell["z"])]
        try:
            with open(filepath, "w") as f:
                f.write("\n".join(lines) + "\n")
            return True
        except FileNotFoundError:
            warn("File error!")
            return False

    def to_XYZ(self, filepath, supercell=False):
        """Write the Kirkland-style XYZ file MSA reads (namsa/supercell.py:
        301-319): '#', a line with the cell extents, '  Z  x  y  z  occ  DW'
        rows, and '-1'. The extents line repeats the x maximum in the y place,
        as the reference writes it (MSA skips that line)."""
        if supercell:
            cell = self.supercell_sites
            x, y, z = cell["x"], cell["y"], cell["z"]
            Z, occ, dw = cell["atomic_number"], cell["occ"], cell["DW"]
        else:
            cell = self.unit_cell_positions
            x, y, z = self.xyz_positions["x"], self.xyz_positions["y"], self.xyz_positions["z"]
            Z, occ, dw = cell["atomic_number"], cell["occ"], cell["DW"]
        lines = ["#", "      %2.4f %2.4f %2.4f" % (x.max(0), x.max(0), z.max(0))]
        lines += ["  %d  %2.4f  %2.4f  %2.4f  %2.4f  %2.4f" % r for r in zip(Z, x, y, z, occ, dw)]
        lines.append("-1")
        try:
            with open(filepath, "w") as f:
                f.write("\n".join(lines) + "\n")
            return True
        except FileNotFoundError:
            warn("File error!")
            return False
