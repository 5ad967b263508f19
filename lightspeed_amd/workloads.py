"""Seeded synthetic inputs for the J/K build: water clusters, an all-trans alkane, densities.

* Water: O-H 0.96 A, H-O-H 104.5 deg, orientation from a seeded uniform random quaternion
  (Shoemake 1992), O atoms uniform in an L^3 cube (angstrom) with rejection sampling to a minimum
  O-O distance of 2.8 A.
* Alkane CnH(2n+2): all-trans zig-zag, C-C 1.54 A, C-H 1.09 A, tetrahedral angles.
* Density: D = C C^T with C the Q factor of a seeded N(0,1) nao x nocc matrix (a symmetric
  projector with |D_ij| <= 1, dense and non-decaying).
"""
import math

import numpy as np

from .loaders import Molecule, Atom, BOHR_PER_ANG


def _random_rotation(rng):
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster_xyz(nwater, box, seed, roh=0.96, angle=104.5, dmin=2.8):
    """(3*nwater, 3) coordinates in angstrom, atom order O, H, H per water."""
    rng = np.random.default_rng(seed)
    half = math.radians(angle) / 2
    local = np.array([[0.0, 0.0, 0.0],
                      [roh * math.sin(half), 0.0, roh * math.cos(half)],
                      [-roh * math.sin(half), 0.0, roh * math.cos(half)]])
    Os = []
    tries = 0
    while len(Os) < nwater:
        tries += 1
        if tries > 2000000:
            raise RuntimeError("water_cluster_xyz: could not place %d waters in a %.1f A cube" % (nwater, box))
        p = rng.random(3) * box
        if all(np.linalg.norm(p - q) >= dmin for q in Os):
            Os.append(p)
    out = []
    for p in Os:
        R = _random_rotation(rng)
        out.extend(p + local @ R.T)
    return np.array(out)


def water_cluster(nwater, box, seed):
    xyz = water_cluster_xyz(nwater, box, seed)
    syms = ["O", "H", "H"] * nwater
    atoms = [Atom(s, s, 8 if s == "O" else 1, x * BOHR_PER_ANG, y * BOHR_PER_ANG, z * BOHR_PER_ANG,
                  8.0 if s == "O" else 1.0, i) for i, (s, (x, y, z)) in enumerate(zip(syms, xyz))]
    return Molecule("h2o_%d" % nwater, atoms)


def alkane(ncarbon, rcc=1.54, rch=1.09):
    """All-trans CnH(2n+2) (angstrom geometry converted to bohr)."""
    theta = math.acos(-1.0 / 3.0)  # tetrahedral
    dx = rcc * math.sin(theta / 2)
    dy = rcc * math.cos(theta / 2)
    C = [np.array([i * dx, (i % 2) * dy, 0.0]) for i in range(ncarbon)]
    atoms = [("C", c) for c in C]
    # two H per carbon out of the chain plane, along the local bisector; terminal carbons get a third
    for i, c in enumerate(C):
        nb = [C[j] for j in (i - 1, i + 1) if 0 <= j < ncarbon]
        bis = -sum((n - c) / np.linalg.norm(n - c) for n in nb)
        bis /= np.linalg.norm(bis)
        zax = np.array([0.0, 0.0, 1.0])
        half = theta / 2
        for s in (+1, -1):
            d = math.cos(half) * bis + s * math.sin(half) * zax
            atoms.append(("H", c + rch * d / np.linalg.norm(d)))
        if len(nb) == 1:
            d = (c - nb[0]) / np.linalg.norm(c - nb[0])
            # terminal H: along the C-C direction rotated into the plane to keep tetrahedral angles
            perp = bis - np.dot(bis, d) * d
            perp /= np.linalg.norm(perp)
            v = math.cos(math.pi - theta) * d - math.sin(math.pi - theta) * perp
            atoms.append(("H", c + rch * v / np.linalg.norm(v)))
    out = [Atom(s, s, 6 if s == "C" else 1, p[0] * BOHR_PER_ANG, p[1] * BOHR_PER_ANG, p[2] * BOHR_PER_ANG,
                6.0 if s == "C" else 1.0, i) for i, (s, p) in enumerate(atoms)]
    return Molecule("c%dh%d" % (ncarbon, 2 * ncarbon + 2), out)


def random_density(nao, nocc, seed):
    """D = C C^T with C (nao x nocc) orthonormal columns from QR of a seeded N(0,1) matrix."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nao, nocc))
    Q, _ = np.linalg.qr(A)
    return Q @ Q.T


def water_cluster_box(n):
    """Cube edge (angstrom) of a seeded (H2O)_n of the batched workload: 10 (n / 10)^(1/3), at least 4."""
    return max(4.0, 10.0 * (n / 10.0) ** (1.0 / 3.0))


def water_batch(nmol=256, seed=3, nmax=16):
    """The batched workload: nmol independent water clusters (H2O)_n with n drawn uniformly from
    1..nmax (seeded), molecule i in a cube of edge water_cluster_box(n) (cluster seed 5000 + i),
    nocc = 5 n, density seed 7000 + i.  Returns a list of (molecule, nocc, density seed)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(1, nmax + 1, size=nmol)
    return [(water_cluster(int(n), water_cluster_box(int(n)), 5000 + i), 5 * int(n), 7000 + i)
            for i, n in enumerate(ns)]


# Named workloads: (molecule builder, basis, nocc, box) -- sizes quoted in the docstrings.
def config(name):
    """Return (molecule, basis_name, nocc) for a named workload.

    h2o10_631g   : (H2O)10, 10 A cube, seed 10, 6-31G   (130 functions, nocc 50)
    c20h42_ccpvdz: C20H42 all-trans, cc-pVDZ spherical   (490 functions, nocc 81)
    h2o64_631gss : (H2O)64, 20 A cube, seed 64, 6-31G**  (1600 functions, nocc 320)
    h2o256_631gs : (H2O)256, 32 A cube, seed 256, 6-31G* (4864 functions, nocc 1280)
    """
    if name == "h2o10_631g":
        return water_cluster(10, 10.0, 10), "6-31g", 50
    if name == "c20h42_ccpvdz":
        return alkane(20), "cc-pvdz", 81
    if name == "h2o64_631gss":
        return water_cluster(64, 20.0, 64), "6-31gss", 320
    if name == "h2o256_631gs":
        return water_cluster(256, 32.0, 256), "6-31gs", 1280
    raise KeyError(name)
