"""Molecule/XYZ and GBS basis loaders.

Behaviour follows the reference's Python loaders:
  * units: src/python/util.py:1-15 (bohr per angstrom = 1 / 0.52917720859);
  * XYZ parsing: src/python/core.py:46-83 (atom lines, ghost atoms) and 98-161 (header lines);
  * GBS parsing and primitive normalisation: src/python/core.py:265-319 and 349-548
    (SP/SPD shells split into separate S/P/D contracted shells; the primitive coefficient c_k
    already carries the normalisation).

The loaders return plain Python records; `lightspeed_amd.Basis` (the native class) is built from
them by `lightspeed_amd.core`.
"""
import math
import os
import re
from collections import namedtuple

ANG_PER_BOHR = 0.52917720859
BOHR_PER_ANG = 1.0 / ANG_PER_BOHR

BASISDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")

_SYMBOLS = [
    "X", "H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG", "AL", "SI", "P", "S", "CL",
    "AR", "K", "CA", "SC", "TI", "V", "CR", "MN", "FE", "CO", "NI", "CU", "ZN", "GA", "GE", "AS", "SE",
    "BR", "KR",
]
ATOMIC_NUMBERS = {s: i for i, s in enumerate(_SYMBOLS)}

_ATOM_RE = re.compile(r"^\s*([a-z]+)(\d*)\s+(\S+)\s+(\S+)\s+(\S+)\s*$", re.IGNORECASE)
_GHOST_RE = re.compile(r"^\s*(Gh)\(([a-z]+)(\d*)\)\s+(\S+)\s+(\S+)\s+(\S+)\s*$", re.IGNORECASE)

Atom = namedtuple("Atom", "label symbol N x y z Z atomIdx")
PrimRecord = namedtuple(
    "PrimRecord", "c e x y z c0 L is_pure aoIdx cartIdx primIdx shellIdx atomIdx")

_SHELL_L = {k: i for i, k in enumerate("SPDFGHIKLMNOQRTUVWXYZ")}


class Molecule:
    """Atoms with coordinates in bohr (only what the Basis loader and the J/K path need)."""

    def __init__(self, name, atoms, charge=0.0, multiplicity=1.0):
        self.name = name
        self.atoms = list(atoms)
        self.charge = charge
        self.multiplicity = multiplicity

    @property
    def natom(self):
        return len(self.atoms)

    @property
    def xyz(self):
        import numpy as np
        return np.array([[a.x, a.y, a.z] for a in self.atoms], dtype=float).reshape(-1, 3)

    @property
    def xyzZ(self):
        import numpy as np
        return np.array([[a.x, a.y, a.z, a.Z] for a in self.atoms], dtype=float).reshape(-1, 4)

    @staticmethod
    def _atom_from_line(line, scale):
        m = _ATOM_RE.match(line)
        if m:
            sym = m.group(1).upper()
            if sym not in ATOMIC_NUMBERS:
                raise ValueError("XYZ Line has unrecognized atomic symbol:\n%s" % line)
            N = ATOMIC_NUMBERS[sym]
            return Atom(m.group(1) + m.group(2), sym, N, float(m.group(3)) * scale,
                        float(m.group(4)) * scale, float(m.group(5)) * scale, float(N), 0)
        m = _GHOST_RE.match(line)
        if m:
            sym = m.group(2).upper()
            if sym not in ATOMIC_NUMBERS:
                raise ValueError("XYZ Line has unrecognized atomic symbol:\n%s" % line)
            N = ATOMIC_NUMBERS[sym]
            return Atom(m.group(1) + "(" + m.group(2) + m.group(3) + ")", sym, N,
                        float(m.group(4)) * scale, float(m.group(5)) * scale,
                        float(m.group(6)) * scale, 0.0, 0)
        raise ValueError("XYZ Line is malformed:\n%s" % line)

    @staticmethod
    def from_xyz_str(strval, scale=BOHR_PER_ANG, name="mol"):
        lines = [l for l in strval.split("\n") if len(l) and not l.isspace()]
        if not lines:
            raise ValueError("No lines in molecule: %s" % name)
        if re.match(r"^\s*(\d+)\s*$", lines[0]):
            lines = lines[1:]
        Q, M = 0.0, 1.0
        m = re.match(r"^\s*(-?\d+)\s+(\d+)\s*$", lines[0])
        if m:
            Q, M = float(m.group(1)), float(m.group(2))
            lines = lines[1:]
        elif not _ATOM_RE.match(lines[0]) and not _GHOST_RE.match(lines[0]):
            lines = lines[1:]
        atoms = [Molecule._atom_from_line(l, scale) for l in lines]
        atoms = [a._replace(atomIdx=i) for i, a in enumerate(atoms)]
        return Molecule(name, atoms, Q, M)

    @staticmethod
    def from_xyz_file(filename, scale=BOHR_PER_ANG):
        m = re.match(r"^(\S+?)(\.xyz)$", os.path.basename(os.path.normpath(filename)))
        if m is None:
            raise ValueError("XYZ File must be of type (PATH/)name.xyz: %s" % filename)
        with open(filename) as fh:
            return Molecule.from_xyz_str(fh.read(), scale=scale, name=m.group(1))

    @staticmethod
    def concatenate(mols, charge=0.0, multiplicity=1.0):
        atoms = []
        for mol in mols:
            atoms += mol.atoms
        atoms = [a._replace(atomIdx=i) for i, a in enumerate(atoms)]
        return Molecule("-".join(m.name for m in mols), atoms, charge, multiplicity)

    def to_xyz_str(self):
        s = "%d\n%s\n" % (len(self.atoms), self.name)
        for a in self.atoms:
            s += "%-4s %22.15f %22.15f %22.15f\n" % (
                a.label, a.x * ANG_PER_BOHR, a.y * ANG_PER_BOHR, a.z * ANG_PER_BOHR)
        return s


def normalized_coefficients(am, N, ws, es):
    """Primitive coefficients with the shell normalisation folded in (core.py:265-292)."""
    pi32 = math.pow(math.pi, 1.5)
    twoL = math.pow(2, am)
    dfact = 1
    for l in range(1, am + 1):
        dfact *= 2 * l - 1
    K = len(ws)
    cs = [ws[k] * math.sqrt(twoL * math.pow(es[k] + es[k], am + 1.5) / (pi32 * dfact)) for k in range(K)]
    V = 0.0
    for k1 in range(K):
        for k2 in range(K):
            V += math.pow(math.sqrt(4 * es[k1] * es[k2]) / (es[k1] + es[k2]), am + 1.5) * ws[k1] * ws[k2]
    V = math.sqrt(N) * math.pow(V, -0.5)
    return [V * c for c in cs]


def _parse_gbs(lines, name):
    body = [l for l in lines if not re.match(r"^\s*$", l) and not re.match(r"^\s*!", l)]
    if re.match(r"^\s*cartesian\s*$", body[0], re.IGNORECASE):
        spherical = False
    elif re.match(r"^\s*spherical\s*$", body[0], re.IGNORECASE):
        spherical = True
    else:
        raise ValueError("GBS File: Where is the cartesian/spherical line?: %s" % name)
    body = body[1:]
    stars = [0] + [i for i, l in enumerate(body) if re.match(r"^\s*\*\*\*\*\s*$", l)] + [len(body)]
    atom_lines = {}
    for k in range(len(stars) - 1):
        i1, i2 = stars[k] + 1, stars[k + 1]
        if i2 - i1 <= 0:
            continue
        m = re.match(r"^\s*(\S+)\s+(\d+)\s*$", body[i1])
        if m is None:
            raise ValueError("GBS File: Where is the ID V line?: %s" % name)
        if m.group(2) != "0":
            continue
        atom_lines[m.group(1).upper()] = body[i1 + 1:i2]
    return spherical, atom_lines


def _shells_for_atom(lines, spherical, normalize):
    """List of contracted shells [(L, [(c, e, c0), ...]), ...] for one element."""
    shells = []
    i = 0
    while i < len(lines):
        m = re.match(r"^\s*(\S+)\s+(\d+)\s+(\S+)\s*$", lines[i])
        if m is None:
            raise ValueError("GBS File: Where is the L K N line?")
        ID = m.group(1).upper()
        K = int(m.group(2))
        N = float(m.group(3))
        i += 1
        rows = []
        for k in range(K):
            rows.append([float(v) for v in re.sub(r"[Dd]", "E", lines[i + k]).split()])
        i += K
        if ID in ("SP", "SPD"):
            Ls = [0, 1] if ID == "SP" else [0, 1, 2]
        else:
            Ls = [_SHELL_L[ID]]
        es = [r[0] for r in rows]
        for j, L in enumerate(Ls):
            ws = [r[1 + j] for r in rows]
            cs = normalized_coefficients(L, N, ws, es) if normalize else list(ws)
            shells.append((L, list(zip(cs, es, ws))))
    return shells


def basis_records_from_gbs_lines(molecule, lines, name="bas", normalize=True, spherical=None):
    """Dense primitive records for `molecule` (core.py:349-502)."""
    for A, atom in enumerate(molecule.atoms):
        if A != atom.atomIdx:
            raise RuntimeError("Basis::from_gbs_lines: cannot parse basis for sparse molecule")
    sph2, atom_lines = _parse_gbs(lines, name)
    if spherical is None:
        spherical = sph2
    cache = {}
    recs = []
    nao = ncart = nprim = nshell = 0
    for A, atom in enumerate(molecule.atoms):
        if atom.symbol not in cache:
            if atom.symbol not in atom_lines:
                raise KeyError(atom.symbol)
            cache[atom.symbol] = _shells_for_atom(atom_lines[atom.symbol], spherical, normalize)
        for (L, prims) in cache[atom.symbol]:
            pnao = (2 * L + 1) if spherical else (L + 1) * (L + 2) // 2
            pncart = (L + 1) * (L + 2) // 2
            for (c, e, c0) in prims:
                recs.append(PrimRecord(c, e, atom.x, atom.y, atom.z, c0, L, bool(spherical),
                                       nao, ncart, nprim, nshell, A))
                nprim += 1
            nao += pnao
            ncart += pncart
            nshell += 1
    return recs


def resolve_gbs(filename):
    if len(filename) >= 4 and filename[-4:] != ".gbs":
        filename += ".gbs"
    if not os.path.isfile(filename):
        filename = os.path.join(BASISDIR, filename)
    if not os.path.isfile(filename):
        raise ValueError("Could not find GBS file: %s" % filename)
    return filename


def basis_records_from_gbs_file(molecule, filename, normalize=True, spherical=None):
    path = resolve_gbs(filename)
    with open(path) as fh:
        lines = fh.readlines()
    name = re.match(r"^(\S+?)(\.gbs)$", os.path.basename(os.path.normpath(path))).group(1)
    return name, basis_records_from_gbs_lines(molecule, lines, name=name, normalize=normalize,
                                              spherical=spherical)
