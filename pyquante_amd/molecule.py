"""Atoms and molecules with PyQuante's unit semantics, plus the seeded
water-cluster generator used by the benchmark workloads.

Mirrors PyQuante/Molecule.py:35-172 and PyQuante/Atom.py: coordinates are
held in bohr; input given in Angstrom is converted with 1/0.529177249.
"""
import math

import numpy as np

from .constants import ang2bohr

_SYM = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al",
        "Si", "P", "S", "Cl", "Ar"]
sym2no = {s: i for i, s in enumerate(_SYM)}


class Atom(object):
    """One nucleus: atomic number, position (bohr), atom id (Atom.py)."""

    def __init__(self, atno, x, y, z, atid=0):
        self.atno = int(atno)
        self.r = np.array([x, y, z], dtype=np.float64)
        self.atid = atid

    def pos(self):
        return tuple(float(v) for v in self.r)

    def get_nuke_chg(self):
        return self.atno

    def get_nel(self):
        return self.atno

    def dist(self, other):
        return float(np.sqrt(np.sum((self.r - other.r) ** 2)))

    def atuple(self):
        return (self.atno, self.pos())

    def update_coords(self, xyz):
        self.r = np.array(xyz, dtype=np.float64)


class Molecule(object):
    """Molecule(name, atomlist, units='bohr'|'angs', charge=0, multiplicity=1)
    (Molecule.py:35-61).  atomlist holds (atno or symbol, (x, y, z))."""

    def __init__(self, name="molecule", atomlist=(), **opts):
        self.name = name
        self.atoms = []
        units = opts.get("units", "bohr").lower()[:4]
        if units not in ("bohr", "angs"):
            raise ValueError("units must be bohr or angs")
        self.units = units
        for atid, (atno, xyz) in enumerate(atomlist):
            if isinstance(atno, str):
                atno = sym2no[atno]
            if units != "bohr":
                xyz = [v * ang2bohr for v in xyz]
            self.atoms.append(Atom(atno, xyz[0], xyz[1], xyz[2], atid))
        self.charge = int(opts.get("charge", 0))
        self.multiplicity = int(opts.get("multiplicity", 1))

    def __iter__(self):
        return iter(self.atoms)

    def __len__(self):
        return len(self.atoms)

    def __getitem__(self, i):
        return self.atoms[i]

    def get_nel(self):
        return sum(a.get_nel() for a in self.atoms) - self.charge

    def get_enuke(self):
        """Nuclear repulsion, summed over pairs j < i (Molecule.py:164-172)."""
        e = 0.0
        for i, ai in enumerate(self.atoms):
            for aj in self.atoms[:i]:
                e += ai.get_nuke_chg() * aj.get_nuke_chg() / ai.dist(aj)
        return e

    def get_closedopen(self):
        """(nclosed, nopen) from electron count and multiplicity (Molecule.py:174-)."""
        nel = self.get_nel()
        nopen = self.multiplicity - 1
        nclosed, rem = divmod(nel - nopen, 2)
        if rem:
            raise ValueError("inconsistent charge/multiplicity")
        return nclosed, nopen

    def get_alphabeta(self):
        """(nalpha, nbeta) (Molecule.py:199-201)."""
        nclosed, nopen = self.get_closedopen()
        return nclosed + nopen, nclosed

    def coords(self):
        return np.array([a.r for a in self.atoms])

    def update_coords(self, coords):
        coords = np.asarray(coords, dtype=np.float64).reshape(len(self.atoms), 3)
        for a, xyz in zip(self.atoms, coords):
            a.update_coords(xyz)

    def copy_with_coords(self, coords):
        m = Molecule(self.name, [(a.atno, tuple(c)) for a, c in
                                 zip(self.atoms, np.asarray(coords).reshape(-1, 3))],
                     units="bohr", charge=self.charge, multiplicity=self.multiplicity)
        return m


def h2o_sto3g_config():
    """H2O with O at the origin and H at (+-0.7572, 0.5865, 0) Angstrom."""
    return Molecule("h2o", [(8, (0.0, 0.0, 0.0)), (1, (0.7572, 0.5865, 0.0)),
                            (1, (-0.7572, 0.5865, 0.0))], units="angs")


def _random_rotation(rng):
    """Uniform random rotation from a random unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(lattice=(2, 2, 2), seed=0, spacing=3.0, jitter=0.3, nwater=None):
    """Seeded (H2O)_n cluster: rigid waters (r_OH = 0.9572 A, HOH = 104.52 deg),
    each randomly rotated, on a cubic lattice of `spacing` Angstrom with a
    uniform +-`jitter` Angstrom translation per molecule and axis.

    numpy.random.default_rng(seed) (PCG64) drives every draw, in the order
    rotation then jitter per molecule, lattice sites in x-major order.  With
    nwater given, the first nwater sites of the lattice are used.
    """
    rng = np.random.default_rng(seed)
    r = 0.9572
    half = math.radians(104.52) / 2
    body = np.array([[0.0, 0.0, 0.0],
                     [r * math.sin(half), r * math.cos(half), 0.0],
                     [-r * math.sin(half), r * math.cos(half), 0.0]])
    com = (16.0 * body[0] + 1.0 * body[1] + 1.0 * body[2]) / 18.0
    body = body - com
    sites = [(i, j, k) for i in range(lattice[0]) for j in range(lattice[1])
             for k in range(lattice[2])]
    if nwater is not None:
        sites = sites[:nwater]
    atoms = []
    for (i, j, k) in sites:
        R = _random_rotation(rng)
        shift = np.array([i, j, k], dtype=np.float64) * spacing + rng.uniform(-jitter, jitter, 3)
        xyz = body @ R.T + shift
        atoms += [(8, tuple(xyz[0])), (1, tuple(xyz[1])), (1, tuple(xyz[2]))]
    return Molecule("h2o_%d" % len(sites), atoms, units="angs")


def lattice_for(nwater):
    """Smallest near-cubic lattice holding nwater sites (for the batched workload)."""
    for n in range(1, 20):
        for shape in ((n, n, n), (n, n, n + 1), (n, n + 1, n + 1)):
            if shape[0] * shape[1] * shape[2] >= nwater:
                if shape == (n, n, n) or shape[0] * shape[1] * shape[2] >= nwater:
                    return shape
    raise ValueError(nwater)
