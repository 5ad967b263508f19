"""HIP J/K at the full sizes of the three single-molecule workloads: C20H42/cc-pVDZ (spherical,
490 functions), (H2O)64/6-31G** (1600) and (H2O)256/6-31G* (4864).

The CPU oracle cannot run these whole builds in seconds, so they are checked two ways:
  * against the oracle on a bounded sample of the same workload: every `stride`-th bra pair of
    each class (PairList.subsample on the GPU side, OPairList.subsample on the oracle side, the
    same pairs) against the whole ket list, FP64, 1e-10 absolute;
  * through properties of the whole build that hold at any size: J is symmetric; the symmetric
    K path (D24symm: rows in triangle order plus the transposed block) equals the full
    non-symmetric K path; the mixed FP32/FP64 route (thredp 1e-6) stays within 1e-7 of the FP64
    route in every element, and within 1e-6 Eh in E2 = 2 tr(DJ) - tr(DK).
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# workload, bra-pair stride of the oracle sample
CASES = [("c20h42_ccpvdz", 96), ("h2o64_631gss", 96), ("h2o256_631gs", 2048)]
THRE = 1.0e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(ls, name):
    from lightspeed_amd import workloads as W
    mol, bname, nocc = W.config(name)
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = W.random_density(bas.nao, nocc, 7)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    return bas, D, pl


@pytest.mark.parametrize("name,stride", CASES)
def test_sample_against_oracle(product_backend, oracle_backend, name, stride):
    ls = product_backend.ls
    O = oracle_backend.O
    _say("full-size sample %s stride %d" % (name, stride))
    bas, D, pl = _setup(ls, name)
    sub = pl.subsample(stride)
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, THRE)
    osub = opl.subsample(stride)
    assert sub.npair == osub.c.npair
    ew = ([1.0], [-1.0])
    t0 = time.perf_counter()
    Jo = O.coulomb(ew, osub, opl, D, THRE)
    Ko = O.exchange(ew, osub, opl, D, False, THRE)
    _say("  oracle on the sample: %.1f s" % (time.perf_counter() - t0))
    J = product_backend.coulomb(ew, sub, pl, D, THRE)
    K = product_backend.exchange(ew, sub, pl, D, False, THRE)
    dJ, dK = np.abs(J - Jo).max(), np.abs(K - Ko).max()
    _say("  max|dJ| = %.2e  max|dK| = %.2e  (|J|max %.2e, |K|max %.2e)" % (dJ, dK, np.abs(Jo).max(), np.abs(Ko).max()))
    assert np.abs(Jo).max() > 0.0 and np.abs(Ko).max() > 0.0
    assert dJ < 1.0e-10 and dK < 1.0e-10


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_whole_build_properties(product_backend, name):
    ls = product_backend.ls
    _say("full-size properties %s" % name)
    bas, D, pl = _setup(ls, name)
    res = product_backend.res
    ew = ls.Ewald([1.0], [-1.0])
    Dt = ls.Tensor.array(D)
    J = np.array(ls.IntBox.coulomb(res, ew, pl, pl, Dt, THRE, THRE).np)
    K = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, True, THRE, THRE).np)
    Kn = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, False, THRE, THRE).np)
    Jm = np.array(ls.IntBox.coulomb(res, ew, pl, pl, Dt, THRE, 1.0e-6).np)
    Km = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, True, THRE, 1.0e-6).np)
    # the symmetric J kernel (one list object on both sides) against the one-way kernels (a second,
    # equal list as the ket list)
    pl2 = ls.PairList.build_schwarz(bas, bas, True, THRE)
    J2 = np.array(ls.IntBox.coulomb(res, ew, pl, pl2, Dt, THRE, THRE).np)
    dJ2 = np.abs(J - J2).max()
    _say("  |J(pl, pl) - J(pl, pl2)| %.1e" % dJ2)
    assert dJ2 < 1.0e-11
    sJ, sK, dKn = np.abs(J - J.T).max(), np.abs(K - K.T).max(), np.abs(K - Kn).max()
    dJm, dKm = np.abs(Jm - J).max(), np.abs(Km - K).max()
    E2 = 2.0 * np.sum(D * J) - np.sum(D * K)
    E2m = 2.0 * np.sum(D * Jm) - np.sum(D * Km)
    _say("  |J-J^T| %.1e  |K-K^T| %.1e  |K_sym - K_full| %.1e  mixed: |dJ| %.1e |dK| %.1e |dE2| %.1e" % (
        sJ, sK, dKn, dJm, dKm, abs(E2m - E2)))
    assert sJ < 1.0e-11 and sK < 1.0e-11
    assert dKn < 1.0e-10
    assert dJm <= 1.0e-7 and dKm <= 1.0e-7 and abs(E2m - E2) <= 1.0e-6


# The batched workload's molecules: (H2O)_n, n = 1..16, 6-31G*, seeded clusters (O atoms in a cube
# of edge 10 A (n / 10)^(1/3), at least 4 A), D from a seeded random orthonormal C with nocc = 5n.
# Each is checked against the oracle: the whole build up to 120 functions, a bra-pair sample
# (every 8th pair of each class, whole ket list) above.
@pytest.mark.parametrize("n", list(range(1, 17)))
def test_water_clusters_1_to_16(product_backend, oracle_backend, n):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    _say("(H2O)%d 6-31G*" % n)
    mol = W.water_cluster(n, max(4.0, 10.0 * (n / 10.0) ** (1.0 / 3.0)), 1000 + n)
    bas = ls.Basis.from_gbs_file(mol, "6-31gs")
    D = W.random_density(bas.nao, 5 * n, 2000 + n)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, THRE)
    stride = 1 if bas.nao <= 120 else 8
    sub, osub = (pl, opl) if stride == 1 else (pl.subsample(stride), opl.subsample(stride))
    ew = ([1.0], [-1.0])
    Jo = O.coulomb(ew, osub, opl, D, THRE)
    J = product_backend.coulomb(ew, sub, pl, D, THRE)
    if stride == 1:
        Ko = O.exchange(ew, opl, opl, D, True, THRE)
        K = product_backend.exchange(ew, pl, pl, D, True, THRE)
    else:
        Ko = O.exchange(ew, osub, opl, D, False, THRE)
        K = product_backend.exchange(ew, sub, pl, D, False, THRE)
    dJ, dK = np.abs(J - Jo).max(), np.abs(K - Ko).max()
    _say("  nao %d stride %d: max|dJ| = %.2e  max|dK| = %.2e" % (bas.nao, stride, dJ, dK))
    assert dJ < 1.0e-10 and dK < 1.0e-10
