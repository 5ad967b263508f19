"""The reference's J/K regression cases, expressed once for any backend.

Each case reproduces one IntBox::coulomb / IntBox::exchange call of the reference's own tests
and carries its stored result:
  * tests/coverage/prop_int.py:21-109   (prop_h2o_{sto-3g,6-31gs,cc-pvdz}.npz, tol 1e-9; every
    quantity computed twice to test accumulation, stored value x2),
  * tests/coverage/ewald_int.py:20-112  (ewald_h2o_*.npz, 5 operators, tol 1e-12),
  * tests/coverage/mix_basis.py:19-250  (mix_h2o_nh3.npz, 2 fragments, tol 1e-12),
  * tests/coverage/mix_basis.py:253-420 (mix_h2o_nh3_ch4_c2h4.npz, 4 fragments, tol 1e-12).
The data files are copies of the reference's fixtures under tests/golden/reference/.

ref_golden_cases adds J and K of the reference's CPU path built from its own sources
(oracle/ref/Makefile) on seeded water clusters, written by tests/golden/make_ref_golden.py
(ref_jk_*.npz: (H2O)10/6-31G, (H2O)4/cc-pVDZ spherical, (H2O)3/6-31G** Cartesian d).

A backend supplies make_basis(mol, name), make_pairlist(b1, b2, symmetric, thre),
coulomb(ewald, pl12, pl34, D, thresp, J=None) and exchange(ewald, pl12, pl34, D, symm, thresp, K=None).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "golden", "reference")

COULOMB = ([1.0], [-1.0])
EWALD_OPS = {
    "LR": ([1.0], [0.3]),
    "LR2": ([0.7, 0.7, 0.7], [0.2, 0.1, 0.05]),
    "SR": ([-0.7, 0.7], [0.3, -1.0]),
    "SR2": ([-0.7, 0.7], [-1.0, -1.0]),
    "MIX": ([0.8, 0.2], [0.2, -1.0]),
}


def molecule(name):
    from lightspeed_amd.loaders import Molecule
    return Molecule.from_xyz_file(os.path.join(DATA, name + ".xyz"))


def prop_cases(backend, basname, symmetric):
    """(key, value, reference, tol) for prop_int.py with one pair-list symmetry."""
    mol = molecule("h2o")
    bas = backend.make_basis(mol, basname)
    z = np.load(os.path.join(DATA, "prop_h2o_%s.npz" % basname))
    pl = backend.make_pairlist(bas, bas, symmetric, 1.0e-14)
    out = []
    D, D2 = z["D"], z["D2"]
    # computed twice to test accumulation (prop_int.py:68-109); reference value doubled
    J = backend.coulomb(COULOMB, pl, pl, D, 1.0e-14)
    J = backend.coulomb(COULOMB, pl, pl, D, 1.0e-14, J)
    out.append(("J", J, 2.0 * z["J"], 1.0e-9))
    J2 = backend.coulomb(COULOMB, pl, pl, D2, 1.0e-14)
    J2 = backend.coulomb(COULOMB, pl, pl, D2, 1.0e-14, J2)
    out.append(("J2", J2, 2.0 * z["J2"], 1.0e-9))
    K = backend.exchange(COULOMB, pl, pl, D, True, 1.0e-14)
    K = backend.exchange(COULOMB, pl, pl, D, True, 1.0e-14, K)
    out.append(("K", K, 2.0 * z["K"], 1.0e-9))
    K2 = backend.exchange(COULOMB, pl, pl, D2, False, 1.0e-14)
    K2 = backend.exchange(COULOMB, pl, pl, D2, False, 1.0e-14, K2)
    out.append(("K2", K2, 2.0 * z["K2"], 1.0e-9))
    return out


def one_electron_cases(backend, basname, symmetric):
    """S, T, V of prop_int.py:53-65 (each computed twice to test accumulation; tol 1e-9; V from the
    nuclear charges mol.xyzZ with the bare Coulomb operator)."""
    mol = molecule("h2o")
    bas = backend.make_basis(mol, basname)
    z = np.load(os.path.join(DATA, "prop_h2o_%s.npz" % basname))
    pl = backend.make_pairlist(bas, bas, symmetric, 1.0e-14)
    S = backend.overlap(pl)
    S = backend.overlap(pl, S)
    T = backend.kinetic(pl)
    T = backend.kinetic(pl, T)
    V = backend.potential(COULOMB, pl, mol.xyzZ)
    V = backend.potential(COULOMB, pl, mol.xyzZ, V)
    return [("S", S, 2.0 * z["S"], 1.0e-9), ("T", T, 2.0 * z["T"], 1.0e-9), ("V", V, 2.0 * z["V"], 1.0e-9)]


def ewald_cases(backend, basname, symmetric):
    mol = molecule("h2o")
    bas = backend.make_basis(mol, basname)
    z = np.load(os.path.join(DATA, "ewald_h2o_%s.npz" % basname))
    pl = backend.make_pairlist(bas, bas, symmetric, 1.0e-14)
    out = []
    for k, ew in EWALD_OPS.items():
        D, D2 = z["D_" + k], z["D2_" + k]
        out.append(("J_" + k, backend.coulomb(ew, pl, pl, D, 1.0e-14), z["J_" + k], 1.0e-12))
        out.append(("J2_" + k, backend.coulomb(ew, pl, pl, D2, 1.0e-14), z["J2_" + k], 1.0e-12))
        out.append(("K_" + k, backend.exchange(ew, pl, pl, D, True, 1.0e-14), z["K_" + k], 1.0e-12))
        out.append(("K2_" + k, backend.exchange(ew, pl, pl, D2, False, 1.0e-14), z["K2_" + k], 1.0e-12))
    return out


def mix2_cases(backend, bas_a, bas_b):
    """mix_basis.py:run_mix_basis_2 for one (basis_a, basis_b) combination."""
    th = 1.0e-14
    mols = [molecule("h2o"), molecule("nh3")]
    refs = np.load(os.path.join(DATA, "mix_h2o_nh3.npz"))
    b1 = backend.make_basis(mols[0], bas_a)
    b2 = backend.make_basis(mols[1], bas_b)
    N1, N2 = backend.nao(b1), backend.nao(b2)
    p11 = backend.make_pairlist(b1, b1, True, th)
    p12 = backend.make_pairlist(b1, b2, False, th)
    p21 = backend.make_pairlist(b2, b1, False, th)
    p22 = backend.make_pairlist(b2, b2, True, th)
    suf = "_" + bas_a.upper() + "_" + bas_b.upper()
    sl = [slice(0, N1), slice(N1, N1 + N2)]

    def blocks(fn, specs, D):
        M = np.zeros((N1 + N2, N1 + N2))
        for (i, j, x, y) in specs:
            M[sl[i], sl[j]] = fn(x, y, D)
        return M

    J = lambda x, y, D: backend.coulomb(COULOMB, x, y, D, th)
    Kf = lambda x, y, D: backend.exchange(COULOMB, x, y, D, False, th)
    Kt = lambda x, y, D: backend.exchange(COULOMB, x, y, D, True, th)
    out = []
    D12 = refs["D0" + suf][:N1, N1:]
    out.append(("J0", blocks(J, [(0, 0, p11, p12), (0, 1, p12, p12), (1, 0, p21, p12), (1, 1, p22, p12)], D12)))
    out.append(("K0", blocks(Kf, [(0, 0, p11, p12), (0, 1, p11, p22), (1, 0, p21, p12), (1, 1, p21, p22)], D12)))
    D22 = refs["D1" + suf][N1:, N1:]
    out.append(("J1", blocks(J, [(0, 0, p11, p22), (0, 1, p12, p22), (1, 0, p21, p22), (1, 1, p22, p22)], D22)))
    out.append(("K1", blocks(Kt, [(0, 0, p12, p12), (0, 1, p12, p22), (1, 0, p22, p12), (1, 1, p22, p22)], D22)))
    D11 = refs["D2" + suf][:N1, :N1]
    out.append(("J2", blocks(J, [(0, 0, p11, p11), (0, 1, p12, p11), (1, 0, p21, p11), (1, 1, p22, p11)], D11)))
    out.append(("K2", blocks(Kt, [(0, 0, p11, p11), (0, 1, p11, p21), (1, 0, p21, p11), (1, 1, p21, p21)], D11)))
    D21 = refs["D3" + suf][N1:, :N1]
    out.append(("J3", blocks(J, [(0, 0, p11, p21), (0, 1, p12, p21), (1, 0, p21, p21), (1, 1, p22, p21)], D21)))
    out.append(("K3", blocks(Kf, [(0, 0, p12, p11), (0, 1, p12, p21), (1, 0, p22, p11), (1, 1, p22, p21)], D21)))
    return [(k + suf, v, refs[k + suf], 1.0e-12) for (k, v) in out]


def mix4_cases(backend, rot):
    """mix_basis.py:run_mix_basis_4 for basis rotation `rot` (0..3)."""
    th = 1.0e-14
    names = ["h2o", "nh3", "ch4", "c2h4"]
    bl = ["sto-3g", "6-31gs", "cc-pvdz", "3-21g"]
    bl2 = bl + bl
    new = bl2[rot:rot + 4]
    mols = [molecule(n) for n in names]
    refs = np.load(os.path.join(DATA, "mix_h2o_nh3_ch4_c2h4.npz"))
    bas = [backend.make_basis(mols[i], new[i]) for i in range(4)]
    P = {}
    for (i, j, s) in [(0, 0, True), (1, 1, True), (2, 2, True), (0, 1, False), (1, 0, False), (0, 2, False),
                      (2, 0, False), (1, 2, False), (2, 1, False), (2, 3, False), (3, 1, False)]:
        P["%d%d" % (i + 1, j + 1)] = backend.make_pairlist(bas[i], bas[j], s, th)
    suf = "".join("_" + b.upper() for b in new)
    D11 = refs["D11" + suf]
    D12 = refs["D12" + suf]
    c = lambda a, b, D: backend.coulomb(COULOMB, P[a], P[b], D, th)
    e = lambda a, b, D, s: backend.exchange(COULOMB, P[a], P[b], D, s, th)
    out = [
        ("J1111", c("11", "11", D11)), ("J1112", c("12", "11", D11)), ("J1122", c("22", "11", D11)),
        ("J1123", c("23", "11", D11)), ("J1234", c("34", "12", D12)), ("J1233", c("33", "12", D12)),
        ("J1223", c("23", "12", D12)), ("J1222", c("22", "12", D12)), ("J1213", c("13", "12", D12)),
        ("J1212", c("12", "12", D12)),
        ("K1111", e("11", "11", D11, True)), ("K1112", e("11", "21", D11, True)),
        ("K1122", e("21", "21", D11, True)), ("K1123", e("21", "31", D11, True)),
        ("K1234", e("31", "42", D12, False)), ("K1233", e("31", "32", D12, False)),
        ("K1223", e("21", "32", D12, False)), ("K1222", e("21", "22", D12, False)),
        ("K1213", e("11", "32", D12, False)), ("K1212", e("11", "22", D12, False)),
    ]
    return [(k + suf, v, refs[k + suf], 1.0e-12) for (k, v) in out]


MIX2_COMBOS = [(a, b) for a in ["sto-3g", "6-31gs", "cc-pvdz"] for b in ["sto-3g", "6-31gs", "cc-pvdz"] if a != b]
BASES = ["sto-3g", "6-31gs", "cc-pvdz"]

REF_GOLDEN = ["h2o10_631g", "h2o4_ccpvdz", "h2o3_631gss"]


def ref_golden_cases(backend, name, tol):
    """J and K (D24symm) of one ref_jk_<name>.npz problem, symmetric Schwarz list."""
    from lightspeed_amd import workloads
    z = np.load(os.path.join(HERE, "golden", "ref_jk_%s.npz" % name))
    mol = workloads.water_cluster(int(z["nwater"]), float(z["box"]), int(z["seed"]))
    b = backend.make_basis(mol, str(z["basis"]))
    pl = backend.make_pairlist(b, b, True, float(z["threpq"]))
    thresp = float(z["thresp"])
    yield name + ":J", backend.coulomb(COULOMB, pl, pl, z["D"], thresp), z["J"], tol
    yield name + ":K", backend.exchange(COULOMB, pl, pl, z["D"], True, thresp), z["K"], tol


class OracleBackend:
    """The CPU oracle (test infrastructure)."""

    def __init__(self):
        import sys
        sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
        import oracle
        self.O = oracle

    def make_basis(self, mol, name):
        from lightspeed_amd import loaders

        class _B:
            pass
        b = _B()
        b.primitives = loaders.basis_records_from_gbs_file(mol, name)[1]
        return self.O.OBasis(b)

    def nao(self, b):
        return b.nao

    def make_pairlist(self, b1, b2, symmetric, thre):
        return self.O.OPairList(b1, b2, symmetric, thre)

    def coulomb(self, ew, pl12, pl34, D, thresp, J=None):
        return self.O.coulomb(ew, pl12, pl34, D, thresp, J)

    def exchange(self, ew, pl12, pl34, D, symm, thresp, K=None):
        return self.O.exchange(ew, pl12, pl34, D, symm, thresp, K)

    def overlap(self, pl, S=None):
        return self.O.overlap(pl, S)

    def kinetic(self, pl, T=None):
        return self.O.kinetic(pl, T)

    def potential(self, ew, pl, xyzc, V=None):
        return self.O.potential(ew, pl, xyzc, V)


class ProductBackend:
    """lightspeed_amd's IntBox (HIP path)."""

    def __init__(self):
        import lightspeed_amd as ls
        self.ls = ls
        self.res = ls.ResourceList.build()

    def make_basis(self, mol, name):
        return self.ls.Basis.from_gbs_file(mol, name)

    def nao(self, b):
        return b.nao

    def make_pairlist(self, b1, b2, symmetric, thre):
        return self.ls.PairList.build_schwarz(b1, b2, symmetric, thre)

    def coulomb(self, ew, pl12, pl34, D, thresp, J=None):
        ls = self.ls
        out = ls.IntBox.coulomb(self.res, ls.Ewald(*ew), pl12, pl34, ls.Tensor.array(np.ascontiguousarray(D)),
                                thresp, thresp, None if J is None else ls.Tensor.array(J))
        return np.array(out.np)

    def exchange(self, ew, pl12, pl34, D, symm, thresp, K=None):
        ls = self.ls
        out = ls.IntBox.exchange(self.res, ls.Ewald(*ew), pl12, pl34, ls.Tensor.array(np.ascontiguousarray(D)),
                                 symm, thresp, thresp, None if K is None else ls.Tensor.array(K))
        return np.array(out.np)

    def overlap(self, pl, S=None):
        ls = self.ls
        return np.array(ls.IntBox.overlap(self.res, pl, None if S is None else ls.Tensor.array(S)).np)

    def kinetic(self, pl, T=None):
        ls = self.ls
        return np.array(ls.IntBox.kinetic(self.res, pl, None if T is None else ls.Tensor.array(T)).np)

    def potential(self, ew, pl, xyzc, V=None):
        ls = self.ls
        out = ls.IntBox.potential(self.res, ls.Ewald(*ew), pl, ls.Tensor.array(np.ascontiguousarray(xyzc)),
                                  None if V is None else ls.Tensor.array(V))
        return np.array(out.np)
