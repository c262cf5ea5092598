"""Seeded synthetic structures in the reference's supercell record layout.

The reference feeds MSA a structured array with fields atomic_number, x, y,
z, occ, DW (namsa/supercell.py:446-450, namsa/msa.py:76-99).  The generators
here produce the same record layout for the benchmark workloads:

  si111_supercell        diamond-cubic Si along [111], orthogonal cell
                         [1-10] x [11-2] x [111], tiled laterally and cropped
                         to a thickness
  random_amorphous_crystal  Z drawn from a set, a jittered simple-cubic
                         lattice in the lower half and random sequential
                         addition with a minimum spacing in the upper half
  random_crystal, multi_material_specs   seeded cubic / tetragonal /
                         hexagonal crystals for uneven multi-material batches
  polycrystal            Voronoi grains of Si-, GaAs- and Au-like lattices in
                         random 3D orientations
"""
from __future__ import annotations

import numpy as np

SUPERCELL_DTYPE = [("atom_type", "U2"), ("atomic_number", "i4"), ("x", "f4"), ("y", "f4"),
                   ("z", "f4"), ("occ", "f4"), ("DW", "f4")]
DW_DEFAULT = 0.07  # Angstrom, the value SupercellBuilder always uses (supercell.py:105-109)

_SYMBOLS = {1: "H", 3: "Li", 6: "C", 7: "N", 8: "O", 13: "Al", 14: "Si", 29: "Cu", 31: "Ga",
            33: "As", 79: "Au", 83: "Bi"}


def make_records(xyz, Z, dw=DW_DEFAULT):
    """Pack positions (A,3) and atomic numbers (A,) into supercell records."""
    xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
    Z = np.asarray(Z, dtype=np.int32).reshape(-1)
    rec = np.zeros(len(Z), dtype=SUPERCELL_DTYPE)
    rec["atomic_number"] = Z
    rec["atom_type"] = [_SYMBOLS.get(int(z), "X") for z in Z]
    rec["x"], rec["y"], rec["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    rec["occ"] = 1.0
    rec["DW"] = dw
    return rec


def si111_supercell(a=5.431, reps=(5, 3), thickness=200.0):
    """Diamond Si with [111] along z, lateral reps of the [1-10] x [11-2] cell.

    The orthogonal cell spans a*(1,-1,0), a*(1,1,-2), a*(1,1,1) (6 cubic cells,
    48 atoms); it is tiled reps[0] x reps[1] laterally and enough times along
    z, then cropped to 0 <= z < thickness.
    """
    basis_fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    basis = np.concatenate([basis_fcc, basis_fcc + 0.25])  # diamond: 8 atoms / cubic cell
    u = np.array([1, -1, 0]) / np.sqrt(2)
    v = np.array([1, 1, -2]) / np.sqrt(6)
    w = np.array([1, 1, 1]) / np.sqrt(3)
    L = np.array([np.sqrt(2), np.sqrt(6), np.sqrt(3)]) * a
    rng = np.arange(-3, 4)
    cells = np.array(np.meshgrid(rng, rng, rng, indexing="ij")).reshape(3, -1).T
    pts = (cells[:, None, :] + basis[None]).reshape(-1, 3) * a
    rot = np.stack([pts @ u, pts @ v, pts @ w], axis=1)
    tol = 1e-6
    keep = np.all((rot >= -tol) & (rot < L - tol), axis=1)
    cell = np.unique(np.round(rot[keep], 6), axis=0)
    assert len(cell) == 48, len(cell)
    nz = int(np.ceil(thickness / L[2])) + 1
    shifts = np.array(np.meshgrid(np.arange(reps[0]), np.arange(reps[1]), np.arange(nz),
                                  indexing="ij")).reshape(3, -1).T * L
    xyz = (shifts[:, None, :] + cell[None]).reshape(-1, 3)
    xyz = xyz[xyz[:, 2] < thickness]
    order = np.lexsort((xyz[:, 0], xyz[:, 1], xyz[:, 2]))
    return make_records(xyz[order], np.full(len(order), 14))


def _rsa(rng, n, lo, hi, dmin, existing):
    """Random sequential addition of n points in the box [lo, hi) at spacing >= dmin."""
    cell = dmin
    dims = np.ceil((hi - lo) / cell).astype(int) + 1
    grid = {}

    def key(p):
        return tuple(((p - lo) // cell).astype(int))

    for p in existing:
        grid.setdefault(key(p), []).append(p)
    out = []
    attempts = 0
    while len(out) < n:
        attempts += 1
        if attempts > 200 * n + 10000:
            raise RuntimeError("random sequential addition did not converge")
        p = lo + rng.random(3) * (hi - lo)
        k = key(p)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dzz in (-1, 0, 1):
                    for q in grid.get((k[0] + dx, k[1] + dy, k[2] + dzz), ()):
                        if np.sum((p - q) ** 2) < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(k, []).append(p)
            out.append(p)
    del dims
    return np.array(out).reshape(-1, 3)


def random_amorphous_crystal(seed=2, lateral=64.0, thickness=300.0, density=0.06,
                             zset=(6, 14, 29, 79), dmin=1.2, jitter=0.1):
    """Seeded amorphous-plus-crystal block.

    Lower half (z < thickness/2): simple-cubic-like lattice with
    n_xy x n_xy x n_z sites at the target density, each jittered uniformly by
    +-jitter Angstrom.  Upper half: random sequential addition at the same
    density with spacing >= dmin (checked against the lattice half too).  Z is
    uniform over zset.  One upper-half atom is pinned at z = thickness so the
    extent is exactly `thickness` (the slice count is floor(extent/dz)).
    """
    rng = np.random.default_rng(seed)
    half = thickness / 2
    n_half = int(round(lateral * lateral * half * density))
    est = (n_half / (half / lateral)) ** (1 / 3)
    exact = [n for n in range(max(1, int(est) - 4), int(est) + 5) if n_half % (n * n) == 0]
    n_xy = min(exact, key=lambda n: abs(n - est)) if exact else int(round(est))
    n_z = int(round(n_half / (n_xy * n_xy)))
    ax = (np.arange(n_xy) + 0.5) * lateral / n_xy
    az = np.arange(n_z) * half / n_z
    X, Y, Zc = np.meshgrid(ax, ax, az, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Zc.ravel()], 1)
    lat[:, :2] += rng.uniform(-jitter, jitter, size=(len(lat), 2))
    lat[:, 2] += rng.uniform(0, jitter, size=len(lat))
    lo = np.array([0.0, 0.0, half])
    hi = np.array([lateral, lateral, thickness])
    border = lat[lat[:, 2] > half - dmin]
    top = _rsa(rng, n_half, lo, hi, dmin, border)
    xyz = np.concatenate([lat, top])
    xyz[:, 2] -= xyz[:, 2].min()
    # pin the extent: the topmost atom moves up to z = thickness (only away
    # from every other atom, so the spacing rule still holds)
    xyz[int(np.argmax(xyz[:, 2])), 2] = thickness
    Z = rng.choice(np.asarray(zset), size=len(xyz))
    return make_records(xyz, Z)


# ---------------------------------------------------------------- multi-material batch
_LATTICES = ("cubic", "tetragonal", "hexagonal")


def random_crystal(rng, lateral, thickness, a_range=(3.0, 7.0), n_basis=(1, 8), z_range=(3, 83),
                   min_sep=1.0):
    """One seeded random crystal tiled to lateral x lateral x thickness Angstrom.

    Lattice: cubic (a, a, a), tetragonal (a, a, c) or hexagonal (a, a, c at
    120 degrees, tiled through its orthohexagonal cell a x a*sqrt(3) x c, two
    lattice points per cell), with a and c in a_range.  The basis has n_basis
    atoms at random fractional positions kept >= min_sep apart (periodic
    images included), each with Z uniform in z_range.
    """
    kind = _LATTICES[int(rng.integers(len(_LATTICES)))]
    a = float(rng.uniform(*a_range))
    c = a if kind == "cubic" else float(rng.uniform(*a_range))
    if kind == "hexagonal":
        cell = np.array([a, a * np.sqrt(3), c])
        centring = np.array([[0, 0, 0], [0.5, 0.5, 0]])
    else:
        cell = np.array([a, a, c])
        centring = np.zeros((1, 3))
    nb = int(rng.integers(n_basis[0], n_basis[1] + 1))
    basis = []
    tries = 0
    while len(basis) < nb and tries < 2000:
        tries += 1
        f = rng.random(3)
        pts = (f + centring) % 1.0
        ok = True
        for q in basis:
            for p in pts:
                d = (p - q + 0.5) % 1.0 - 0.5
                if np.linalg.norm(d * cell) < min_sep:
                    ok = False
        if ok and all(np.linalg.norm(((pts[i] - pts[j] + 0.5) % 1.0 - 0.5) * cell) >= min_sep
                      for i in range(len(pts)) for j in range(i)):
            basis.append(f)
    basis = np.array(basis)
    Zb = rng.integers(z_range[0], z_range[1] + 1, size=len(basis))
    frac = ((basis[:, None, :] + centring[None]) % 1.0).reshape(-1, 3)
    Zc = np.repeat(Zb, len(centring))
    reps = np.ceil(np.array([lateral, lateral, thickness]) / cell).astype(int)
    grid = np.array(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij")).reshape(3, -1).T
    xyz = ((grid[:, None, :] + frac[None]) * cell).reshape(-1, 3)
    Z = np.tile(Zc, len(grid))
    keep = (xyz[:, 0] < lateral) & (xyz[:, 1] < lateral) & (xyz[:, 2] < thickness)
    return make_records(xyz[keep], Z[keep]), {"lattice": kind, "a": a, "c": c, "basis": len(basis)}


def multi_material_specs(seed=3, count=48, grids=(128, 192, 256, 384, 512, 768, 1024),
                         thickness=(50.0, 500.0), dz=(1.0, 3.0), kv=(60.0, 200.0),
                         semi_mrad=(10.0, 30.0), probes=(1, 256), pix=(0.08, 0.14)):
    """Seeded specifications of an uneven multi-material batch.

    Each entry: records, grid N, dz, kV, semi-angle (rad), number of probe
    positions (a square-ish raster) and the crystal description.  The lateral
    size is N * pix with pix uniform in `pix`, so every grid samples its own
    crystal at 0.08-0.14 Angstrom per pixel.
    """
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        N = int(rng.choice(grids))
        lateral = N * float(rng.uniform(*pix))
        T = float(rng.uniform(*thickness))
        rec, info = random_crystal(rng, lateral, T)
        n_probe = int(rng.integers(probes[0], probes[1] + 1))
        gx = int(np.floor(np.sqrt(n_probe)))
        gy = int(np.ceil(n_probe / gx))
        specs.append({"name": f"crystal{i:02d}", "records": rec, "N": N, "dz": float(rng.uniform(*dz)),
                      "kv": float(rng.uniform(*kv)), "semi": float(rng.uniform(*semi_mrad)) * 1e-3,
                      "grid_steps": (gx, gy), "thickness": T, "lateral": lateral, **info})
    return specs


# ---------------------------------------------------------------- polycrystal
# (name, cubic lattice constant A, basis in fractional coordinates, Z per basis site)
_FCC = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
_GRAIN_LATTICES = (
    ("Si", 5.431, np.concatenate([_FCC, _FCC + 0.25]), [14] * 4 + [14] * 4),
    ("GaAs", 5.653, np.concatenate([_FCC, _FCC + 0.25]), [31] * 4 + [33] * 4),
    ("Au", 4.078, _FCC, [79] * 4),
)


def _random_rotation(rng):
    """Uniformly distributed rotation matrix (QR of a Gaussian matrix, sign-fixed)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def polycrystal(seed=5, lateral=128.0, thickness=1000.0, grains=4, slab=50.0, dmin=1.2):
    """Seeded Voronoi polycrystal filling lateral x lateral x thickness A.

    `grains` seeds are uniform in the box; each grain takes one of the
    diamond-Si, zincblende-GaAs and fcc-Au lattices (cycled, so 4 grains use
    all three) in a uniformly random 3D orientation, and keeps the lattice
    sites whose nearest seed is its own.  Sites are generated slab by slab
    along z (`slab` A thick) so memory stays bounded.  Across grain
    boundaries, of two atoms closer than `dmin` the later one is dropped
    (within a grain the closest pair is 2.35 A).  One atom is pinned at
    z = thickness so the extent is exact.  Returns (records, info).
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    box = np.array([lateral, lateral, thickness])
    seeds = rng.uniform(0, 1, size=(grains, 3)) * box
    kinds = [_GRAIN_LATTICES[g % len(_GRAIN_LATTICES)] for g in range(grains)]
    rots = [_random_rotation(rng) for _ in range(grains)]
    origins = rng.uniform(0, 1, size=(grains, 3))
    xyz_all, Z_all = [], []
    for z0 in np.arange(0.0, thickness, slab):
        z1 = min(thickness, z0 + slab)
        lo = np.array([0.0, 0.0, z0])
        hi = np.array([lateral, lateral, z1])
        corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        for g in range(grains):
            name, a, basis, Zb = kinds[g]
            Rm = rots[g]
            # lattice frame: p_lab = Rm @ (a (n + b + origin))  =>  n ~ Rm^T p_lab / a
            f = corners @ Rm / a - origins[g]
            nlo = np.floor(f.min(0)).astype(int) - 1
            nhi = np.ceil(f.max(0)).astype(int) + 1
            grid = np.stack(np.meshgrid(*[np.arange(nlo[i], nhi[i]) for i in range(3)], indexing="ij"),
                            -1).reshape(-1, 1, 3)
            frac = (grid + basis[None] + origins[g]).reshape(-1, 3)
            lab = (frac * a) @ Rm.T
            inside = np.all((lab >= lo) & (lab < hi), axis=1)
            lab = lab[inside]
            zs = np.tile(np.asarray(Zb), len(grid))[inside]
            d2 = ((lab[:, None, :] - seeds[None]) ** 2).sum(-1)
            mine = np.argmin(d2, axis=1) == g
            xyz_all.append(lab[mine])
            Z_all.append(zs[mine])
    xyz = np.concatenate(xyz_all)
    Z = np.concatenate(Z_all)
    pairs = cKDTree(xyz).query_pairs(r=dmin, output_type="ndarray")
    if len(pairs):
        drop = np.zeros(len(xyz), bool)
        for i, j in pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]:
            if not drop[i]:
                drop[j] = True
        xyz, Z = xyz[~drop], Z[~drop]
    top = int(np.argmax(xyz[:, 2]))
    xyz[top, 2] = thickness
    info = {"grains": [k[0] for k in kinds], "seeds": seeds.tolist(), "atoms": int(len(Z))}
    return make_records(xyz, Z), info
