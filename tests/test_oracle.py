"""The CPU oracle against the reference's own stored J/K results (CPU only).

These pin the oracle (oracle/jk_oracle.c) before it is used to judge the HIP path: a dropped
term, a wrong sign, a transposed operand or a wrong screening predicate shows up as a
mismatch against fixtures the reference computed itself.
"""
import numpy as np
import pytest

import fixture_cases as FC


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_prop_fixtures(oracle_backend, basname, symmetric):
    for key, val, ref, tol in FC.prop_cases(oracle_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_one_electron_fixtures(oracle_backend, basname, symmetric):
    for key, val, ref, tol in FC.one_electron_cases(oracle_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_ewald_fixtures(oracle_backend, basname, symmetric):
    for key, val, ref, tol in FC.ewald_cases(oracle_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("combo", FC.MIX2_COMBOS)
def test_mix2_fixtures(oracle_backend, combo):
    for key, val, ref, tol in FC.mix2_cases(oracle_backend, *combo):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("rot", range(4))
def test_mix4_fixtures(oracle_backend, rot):
    for key, val, ref, tol in FC.mix4_cases(oracle_backend, rot):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("name", FC.REF_GOLDEN)
def test_reference_build_golden(oracle_backend, name):
    """J/K of the reference's CPU path compiled from its own sources (FP64 both sides)."""
    for key, val, ref, tol in FC.ref_golden_cases(oracle_backend, name, 1.0e-12):
        assert np.abs(val - ref).max() < tol, key


def test_reference_build_live():
    """Where the reference build exists (a container with the reference sources), the oracle
    equals it on a fresh seeded problem with a non-symmetric K call as well."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "ref"))
    import refjk
    if not refjk.available():
        pytest.skip("reference build (oracle/_ref/ref_jk) not present")
    import oracle as O
    mol = W.water_cluster(5, 7.5, 123)
    bas = ls.Basis.from_gbs_file(mol, "6-31gs")
    D = W.random_density(bas.nao, 25, 9)
    for symm in (True, False):
        J, K, _ = refjk.run(bas, D, 1e-13, 1e-12, D24symm=symm, J=symm)
        ob = O.OBasis(bas)
        pl = O.OPairList(ob, ob, True, 1e-13)
        if symm:
            assert np.abs(O.coulomb(FC.COULOMB, pl, pl, D, 1e-12) - J).max() < 1e-12
        assert np.abs(O.exchange(FC.COULOMB, pl, pl, D, symm, 1e-12) - K).max() < 1e-12
    # the bounded sample used for the CPU baseline: every 3rd bra pair of each class
    J, K, _ = refjk.run(bas, D, 1e-13, 1e-12, stride=3)
    ob = O.OBasis(bas)
    pl = O.OPairList(ob, ob, True, 1e-13)
    sub = pl.subsample(3)
    assert np.abs(O.coulomb(FC.COULOMB, sub, pl, D, 1e-12) - J).max() < 1e-12
    assert np.abs(O.exchange(FC.COULOMB, sub, pl, D, True, 1e-12) - K).max() < 1e-12


# ---------------------------------------------------------------- closed forms and invariants

import math  # noqa: E402

import lightspeed_amd as ls  # noqa: E402
from lightspeed_amd import workloads as W  # noqa: E402


def _s_basis(centers, exps):
    P = ls.Primitive
    prims = [P(1.0, e, x, y, z, 1.0, 0, False, i, i, i, i, i) for i, ((x, y, z), e) in enumerate(zip(centers, exps))]
    return ls.Basis("s", prims)


def _ssss(A, a, B, b, C, c, D, d):
    """Textbook (ss|ss) for unnormalised primitives: 2 pi^{5/2}/(pq sqrt(p+q)) K_AB K_CD F0(T)."""
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    Kab = math.exp(-a * b / p * np.sum((A - B) ** 2))
    Kcd = math.exp(-c * d / q * np.sum((C - D) ** 2))
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    F0 = 1.0 if T < 1e-14 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0


def test_ssss_closed_form(oracle_backend):
    O = oracle_backend.O
    rng = np.random.default_rng(4)
    cen = rng.uniform(-2, 2, (4, 3))
    ex = rng.uniform(0.3, 3.0, 4)
    bas = O.OBasis(_s_basis(cen, ex))
    pl = O.OPairList(bas, bas, False, 1e-30)
    for (i, j, k, l) in [(0, 1, 2, 3), (0, 0, 1, 1), (3, 2, 0, 1), (1, 1, 1, 1)]:
        D = np.zeros((4, 4))
        D[k, l] = 1.0
        J = O.coulomb(FC.COULOMB, pl, pl, D, 1e-30)
        ref = _ssss(cen[i], ex[i], cen[j], ex[j], cen[k], ex[k], cen[l], ex[l])
        assert abs(J[i, j] - ref) < 1e-13 * max(1, abs(ref))
        D = np.zeros((4, 4))
        D[j, l] = 1.0
        K = O.exchange(FC.COULOMB, pl, pl, D, False, 1e-30)
        assert abs(K[i, k] - ref) < 1e-13 * max(1, abs(ref))


def test_linearity_symmetry_translation(oracle_backend):
    O = oracle_backend.O
    mol = W.water_cluster(3, 6.0, 9)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    ob = O.OBasis(bas)
    pl = O.OPairList(ob, ob, True, 1e-14)
    rng = np.random.default_rng(1)
    D1 = rng.standard_normal((bas.nao, bas.nao))
    D1 = D1 + D1.T
    D2 = W.random_density(bas.nao, 7, 2)
    J1, J2 = O.coulomb(FC.COULOMB, pl, pl, D1, 1e-20), O.coulomb(FC.COULOMB, pl, pl, D2, 1e-20)
    J12 = O.coulomb(FC.COULOMB, pl, pl, D1 + 2 * D2, 1e-20)
    assert np.abs(J12 - (J1 + 2 * J2)).max() < 1e-11
    assert np.abs(J1 - J1.T).max() < 1e-12
    K2 = O.exchange(FC.COULOMB, pl, pl, D2, True, 1e-20)
    assert np.abs(K2 - K2.T).max() < 1e-12
    # rigid translation leaves J and K unchanged
    shift = np.array([1.7, -0.4, 2.9])
    prims = [ls.Primitive(p.c, p.e, p.x + shift[0], p.y + shift[1], p.z + shift[2], p.c0, p.L, p.is_pure, p.aoIdx,
                          p.cartIdx, p.primIdx, p.shellIdx, p.atomIdx) for p in bas.primitives]
    obs = O.OBasis(ls.Basis("shifted", prims))
    pls = O.OPairList(obs, obs, True, 1e-14)
    assert np.abs(O.coulomb(FC.COULOMB, pls, pls, D2, 1e-20) - J2).max() < 1e-11
    assert np.abs(O.exchange(FC.COULOMB, pls, pls, D2, True, 1e-20) - K2).max() < 1e-11
