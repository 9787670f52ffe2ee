"""Seeded synthetic systems of the engine's workloads (all in LAMMPS real units: Angstrom, g/mol, Angstrom/fs).

* water_box: liquid water at 1.0 g/cm^3, molecules on a k x k x k lattice with random orientations (O-H 0.96 A, H-O-H 104.5 deg):
  216 molecules (648 atoms, 18.62 A cube) and 8000 molecules (24,000 atoms, 62.09 A cube) are the two water workloads.
* random_carbon: N carbon atoms placed uniformly at random in a cube of the given density, overlaps closer than dmin resolved by
  re-placing one atom of each close pair (random sequential addition; the method of the reference's
  examples/graphitization/make_box.py:16-27): 30,000 atoms at 1.0 g/cm^3 is an 84.265 A cube.
* ch4_o2_gas: CH4 : O2 = 1 : 2 molecules at random positions and orientations, >= 2.5 A between molecules.
* maxwell_boltzmann: velocities at temperature T from a seed, zero total momentum.
Element order of every system is (C, H, O) -> LAMMPS types 1, 2, 3.
"""
import numpy as np

AMU_G = 1.66053906660e-24
MASS = {"C": 12.0107, "H": 1.00794, "O": 15.9994}
ELEMENTS = ("C", "H", "O")
BOLTZ = 0.0019872067          # kcal/mol/K
MVV2E = 48.88821291 ** 2      # (g/mol) (A/fs)^2 -> kcal/mol


def _rotations(rng, n):
    """n uniformly random rotation matrices (from unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


class System:
    def __init__(self, x, types, box_lo, box_hi, name):
        self.x = np.asarray(x, dtype=np.float64)
        self.types = np.asarray(types, dtype=np.int32)      # LAMMPS types: 1 C, 2 H, 3 O
        self.tags = np.arange(1, len(self.types) + 1, dtype=np.int64)
        self.lo = np.asarray(box_lo, dtype=np.float64)
        self.hi = np.asarray(box_hi, dtype=np.float64)
        self.periodic = (True, True, True)
        self.name = name

    @property
    def n(self):
        return len(self.types)

    def box12(self):
        L = self.hi - self.lo
        return [self.lo[0], self.lo[1], self.lo[2], L[0], 0, 0, 0, L[1], 0, 0, 0, L[2]]

    def masses(self):
        return np.array([MASS[ELEMENTS[t - 1]] for t in self.types])


def water_box(nside=6, seed=1, roh=0.96, hoh=104.5, density=1.0):
    nmol = nside ** 3
    mmol = MASS["O"] + 2 * MASS["H"]
    L = (nmol * mmol * AMU_G / density * 1e24) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    a = np.radians(hoh / 2)
    mol = np.array([[0, 0, 0], [roh * np.sin(a), 0, roh * np.cos(a)], [-roh * np.sin(a), 0, roh * np.cos(a)]])
    mol -= (MASS["O"] * mol[0] + MASS["H"] * (mol[1] + mol[2])) / mmol   # rotate about the centre of mass
    R = _rotations(rng, nmol)
    g = (np.arange(nside) + 0.5) * (L / nside)
    centres = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = (centres[:, None, :] + np.einsum("mij,aj->mai", R, mol)).reshape(-1, 3)
    types = np.tile([3, 2, 2], nmol)
    return System(x, types, [0, 0, 0], [L, L, L], "water_%d" % (3 * nmol))


def random_carbon(n=30000, density=1.0, dmin=1.2, seed=1):
    L = (n * MASS["C"] * AMU_G / density * 1e24) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)) * L
    from scipy.spatial import cKDTree
    for _ in range(1000):
        pairs = cKDTree(x, boxsize=L).query_pairs(dmin, output_type="ndarray")
        if len(pairs) == 0:
            break
        bad = np.unique(pairs[:, 1])
        x[bad] = rng.random((len(bad), 3)) * L
    else:
        raise RuntimeError("could not remove all overlaps")
    return System(x, np.ones(n, dtype=np.int32), [0, 0, 0], [L, L, L], "carbon_%d" % n)


def ch4_o2_gas(n_target=600, density=0.3, seed=1, dmin=2.5):
    """CH4:O2 = 1:2 at `density`; about n_target atoms (a unit = 1 CH4 + 2 O2 = 9 atoms)."""
    units = max(1, int(round(n_target / 9)))
    m = units * (MASS["C"] + 4 * MASS["H"] + 4 * MASS["O"])
    L = (m * AMU_G / density * 1e24) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    t = 1.087 / np.sqrt(3)
    ch4 = np.array([[0, 0, 0], [t, t, t], [t, -t, -t], [-t, t, -t], [-t, -t, t]])
    o2 = np.array([[0, 0, -0.604], [0, 0, 0.604]])
    mols = [(ch4, [1, 2, 2, 2, 2])] * units + [(o2, [3, 3])] * (2 * units)
    order = rng.permutation(len(mols))
    rc = dmin + 1.2                      # minimum distance between molecule centres
    nc = max(1, int(L // rc))            # cells at least rc wide: a clash can only come from the 27 surrounding cells
    cells = {}
    xs, ts = [], []
    for k in order:
        geom, ty = mols[k]
        for _ in range(10000):
            c = rng.random(3) * L
            cc = np.minimum((c / L * nc).astype(int), nc - 1)
            clash = False
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for o in cells.get(((cc[0] + dx) % nc, (cc[1] + dy) % nc, (cc[2] + dz) % nc), ()):
                            d = o - c
                            d -= L * np.round(d / L)
                            if d @ d < rc * rc:
                                clash = True
                                break
                        if clash:
                            break
                    if clash:
                        break
                if clash:
                    break
            if not clash:
                break
        else:
            raise RuntimeError("gas packing failed")
        cells.setdefault(tuple(cc), []).append(c)
        R = _rotations(rng, 1)[0]
        xs.append(c + geom @ R.T)
        ts += ty
    return System(np.concatenate(xs), np.array(ts, dtype=np.int32), [0, 0, 0], [L, L, L], "gas_%d" % len(ts))


def maxwell_boltzmann(masses, T, seed=1):
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, dtype=np.float64)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(BOLTZ * T / (m * MVV2E))[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


def load_ffield(path):
    """The text of a ReaxFF force-field file (the engine parses it: rxa_pair_coeff)."""
    with open(path) as fh:
        return fh.read()
