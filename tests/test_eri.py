"""IntBox::eriJ (reference mo_eri.cpp:10-157): MO-basis ERIs through J builds.

CPU: the oracle's restatement against the AO J it is built from (sum over a diagonal MO pair
reproduces the J build with the corresponding density) and the 8-fold permutational symmetry of
(ij|kl) for real orbitals.  GPU: the HIP path against the oracle at 1e-10.
"""
import numpy as np
import pytest

import fixture_cases as FC


def _setup(O, ls, basname="cc-pvdz", n=3, seed=3):
    mol = FC.molecule("h2o")
    bas = ls.Basis.from_gbs_file(mol, basname)
    rng = np.random.default_rng(seed)
    C = np.linalg.qr(rng.standard_normal((bas.nao, n)))[0]
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, 1.0e-14)
    return bas, C, opl


def test_eriJ_oracle_identities(oracle_backend):
    import lightspeed_amd as ls
    O = oracle_backend.O
    bas, C, opl = _setup(O, ls)
    I = O.eriJ(FC.COULOMB, opl, opl, C, C, C, C, 1.0e-14, True, True, True)
    # 8-fold symmetry
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(I - I.transpose(perm)).max() < 1e-12
    # sum_k (ij|kk) = C_i^T J[sum_k C_k C_k^T] C_j
    D = C @ C.T
    J = O.coulomb(FC.COULOMB, opl, opl, D, 1.0e-14)
    lhs = np.einsum("ijkk->ij", I)
    assert np.abs(lhs - C.T @ J @ C).max() < 1e-11


@pytest.mark.gpu
def test_eriJ_gpu_against_oracle(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    bas, C, opl = _setup(O, ls, n=4)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-14)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    Ct = ls.Tensor.array(C)
    I = np.array(ls.IntBox.eriJ(res, ew, pl, pl, Ct, Ct, Ct, Ct, 1.0e-14, 1.0e-14).np)
    Io = O.eriJ(FC.COULOMB, opl, opl, C, C, C, C, 1.0e-14, True, True, True)
    assert np.abs(I - Io).max() < 1e-10
    # distinct coefficient tensors: no symmetrisation
    rng = np.random.default_rng(8)
    C2 = rng.standard_normal((bas.nao, 2))
    I2 = np.array(ls.IntBox.eriJ(res, ew, pl, pl, Ct, ls.Tensor.array(C2), Ct, ls.Tensor.array(C2), 1.0e-14,
                                 1.0e-14).np)
    I2o = O.eriJ(FC.COULOMB, opl, opl, C, C2, C, C2, 1.0e-14)
    assert np.abs(I2 - I2o).max() < 1e-10
