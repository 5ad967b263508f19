"""HIP path parity (MI355X): IntBox::coulomb / IntBox::exchange through liblsjk against the
reference's stored results and against the CPU oracle.

Tolerances: the reference's own (1e-9 prop, 1e-12 ewald/mix) for the fixtures; 1e-10 absolute
against the FP64 oracle at the (H2O)10/6-31G workload.  Quartet counts (the screening decisions)
must agree exactly with the oracle's.
"""
import sys

import numpy as np
import pytest

import fixture_cases as FC

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_prop_fixtures_gpu(product_backend, basname, symmetric):
    _say("prop %s sym=%s" % (basname, symmetric))
    for key, val, ref, tol in FC.prop_cases(product_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_one_electron_fixtures_gpu(product_backend, basname, symmetric):
    _say("one-electron %s sym=%s" % (basname, symmetric))
    for key, val, ref, tol in FC.one_electron_cases(product_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("basname", FC.BASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_ewald_fixtures_gpu(product_backend, basname, symmetric):
    _say("ewald %s sym=%s" % (basname, symmetric))
    for key, val, ref, tol in FC.ewald_cases(product_backend, basname, symmetric):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("combo", FC.MIX2_COMBOS)
def test_mix2_fixtures_gpu(product_backend, combo):
    _say("mix2 %s" % (combo,))
    for key, val, ref, tol in FC.mix2_cases(product_backend, *combo):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("rot", range(4))
def test_mix4_fixtures_gpu(product_backend, rot):
    _say("mix4 rot=%d" % rot)
    for key, val, ref, tol in FC.mix4_cases(product_backend, rot):
        assert np.abs(val - ref).max() < tol, key


@pytest.mark.parametrize("name", FC.REF_GOLDEN)
def test_reference_build_golden_gpu(product_backend, name):
    """J/K of the reference's CPU path compiled from its own sources, tolerance 1e-11."""
    _say("reference-build golden %s" % name)
    for key, val, ref, tol in FC.ref_golden_cases(product_backend, name, 1.0e-11):
        assert np.abs(val - ref).max() < tol, key


def _water10(ls):
    from lightspeed_amd import workloads as W
    mol, bname, nocc = W.config("h2o10_631g")
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = W.random_density(bas.nao, nocc, 1)
    return mol, bas, D


def test_h2o10_631g_against_oracle(product_backend, oracle_backend):
    """(H2O)10 / 6-31G (130 functions), Schwarz 1e-12, random orthonormal C (nocc 50)."""
    ls = product_backend.ls
    _say("h2o10 6-31G vs oracle")
    mol, bas, D = _water10(ls)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ob = oracle_backend.O.OBasis(bas)
    opl = oracle_backend.O.OPairList(ob, ob, True, 1.0e-12)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    oj_counts = np.zeros(625, np.int64)
    ok_counts = np.zeros(625, np.int64)
    Jo = oracle_backend.O.coulomb(FC.COULOMB, opl, opl, D, 1.0e-12, counts=oj_counts)
    Ko = oracle_backend.O.exchange(FC.COULOMB, opl, opl, D, True, 1.0e-12)
    Ko_ns = oracle_backend.O.exchange(FC.COULOMB, opl, opl, D, False, 1.0e-12, counts=ok_counts)
    ls.set_counting(True)
    try:
        J = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), 1.0e-12, 1.0e-12).np
        jc = ls.last_counts()
        K = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, 1.0e-12, 1.0e-12).np
        K_ns = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), False, 1.0e-12, 1.0e-12).np
        kc = ls.last_counts()
    finally:
        ls.set_counting(False)
    assert np.abs(J - Jo).max() < 1.0e-10
    assert np.abs(K - Ko).max() < 1.0e-10
    assert np.abs(K_ns - Ko_ns).max() < 1.0e-10
    # screening parity: J counts per (L12, L34), K counts per (L1, L2, L3, L4) class
    oj = np.zeros(25, np.int64)
    for L1 in range(5):
        for L2 in range(5):
            for L3 in range(5):
                for L4 in range(5):
                    n = oj_counts[((L1 * 5 + L2) * 5 + L3) * 5 + L4]
                    if n:
                        oj[(L1 + L2) * 5 + (L3 + L4)] += n
    assert np.array_equal(jc[:25].astype(np.int64), oj)
    assert np.array_equal(kc[256:256 + 625].astype(np.int64), ok_counts)


def test_device_pointer_path(product_backend):
    """coulomb_device / exchange_device on HBM-resident buffers equal the host-API results and
    accumulate into their outputs."""
    ls = product_backend.ls
    _say("device pointer path")
    mol, bas, D = _water10(ls)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    J = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), 1.0e-12, 1.0e-12).np
    K = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, 1.0e-12, 1.0e-12).np
    n = bas.ncart
    dD = ls.DeviceArray([n, n])
    dD.upload(D)
    dJ = ls.DeviceArray([n, n])
    dJ.upload(np.ones((n, n)))
    dK = ls.DeviceArray([n, n])
    ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, 1.0e-12, 1.0e-12, dJ.ptr)
    ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, 1.0e-12, 1.0e-12, dK.ptr)
    ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, 1.0e-12, 1.0e-12, dK.ptr)
    assert np.abs(dJ.download() - (J + 1.0)).max() < 1.0e-12
    assert np.abs(dK.download() - 2.0 * K).max() < 1.0e-12


def test_reference_convention_counts(product_backend, oracle_backend):
    """lsjk_count_quartets (the metric's denominator) equals the oracle's evaluated-quartet counts
    per class, for J and for K with and without the D24symm half-sieve."""
    ls = product_backend.ls
    _say("reference-convention counts")
    mol, bas, D = _water10(ls)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ob = oracle_backend.O.OBasis(bas)
    opl = oracle_backend.O.OPairList(ob, ob, True, 1.0e-12)
    cj = np.zeros(625, np.int64)
    ck = np.zeros(625, np.int64)
    ck_ns = np.zeros(625, np.int64)
    oracle_backend.O.coulomb(FC.COULOMB, opl, opl, D, 1.0e-12, counts=cj)
    oracle_backend.O.exchange(FC.COULOMB, opl, opl, D, True, 1.0e-12, counts=ck)
    oracle_backend.O.exchange(FC.COULOMB, opl, opl, D, False, 1.0e-12, counts=ck_ns)
    dD = ls.DeviceArray([bas.ncart, bas.ncart])
    dD.upload(D)
    nj, nk = ls.count_quartets(pl, pl, dD.ptr, dD.ptr, True, 1.0e-12)
    _, nk_ns = ls.count_quartets(pl, pl, 0, dD.ptr, False, 1.0e-12)
    assert np.array_equal(nj.astype(np.int64), cj)
    assert np.array_equal(nk.astype(np.int64), ck)
    assert np.array_equal(nk_ns.astype(np.int64), ck_ns)


# ---------------------------------------------------------------- mixed precision route

def _mixed_problem(ls, name, seed=1):
    from lightspeed_amd import workloads as W
    mol, bname, nocc = W.config(name)
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = W.random_density(bas.nao, nocc, seed)
    return bas, D


def _fold_mirrors(counts):
    """K counters (lsjk_last_counts layout) summed over each pair of mirror classes
    (L1 L2 | L3 L4) / (L3 L4 | L1 L2), keyed by the lower index."""
    out = {}
    for L1 in range(3):
        for L2 in range(3):
            for L3 in range(3):
                for L4 in range(3):
                    c = 256 + ((L1 * 5 + L2) * 5 + L3) * 5 + L4
                    m = 256 + ((L3 * 5 + L4) * 5 + L1) * 5 + L2
                    out[min(c, m)] = out.get(min(c, m), 0) + int(counts[c])
    return np.array([out[k] for k in sorted(out)], np.int64)


def test_mixed_route_h2o10_against_oracle(product_backend, oracle_backend):
    """thredp > thresp: FP64 for est >= thredp, FP32 for thresp <= est < thredp, FP64 accumulation
    (intbox.hpp:736-742).  Against the FP64 oracle: max|dJ|, max|dK| <= 1e-7 at thredp = 1e-6 (the
    accuracy bar of the mixed route); the FP64 and FP32 launches together evaluate exactly the
    quartets of the plain route (per-class counts)."""
    ls = product_backend.ls
    _say("mixed route h2o10 vs oracle")
    bas, D = _mixed_problem(ls, "h2o10_631g")
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ob = oracle_backend.O.OBasis(bas)
    opl = oracle_backend.O.OPairList(ob, ob, True, 1.0e-12)
    Jo = oracle_backend.O.coulomb(FC.COULOMB, opl, opl, D, 1.0e-12)
    Ko = oracle_backend.O.exchange(FC.COULOMB, opl, opl, D, True, 1.0e-12)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    ls.set_counting(True)
    try:
        J0 = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), 1.0e-12, 1.0e-12).np
        cj0 = ls.last_counts().astype(np.int64)
        J = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), 1.0e-12, 1.0e-6).np
        cj = ls.last_counts().astype(np.int64)
        K0 = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, 1.0e-12, 1.0e-12).np
        ck0 = ls.last_counts().astype(np.int64)
        K = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, 1.0e-12, 1.0e-6).np
        ck = ls.last_counts().astype(np.int64)
    finally:
        ls.set_counting(False)
    # the symmetric K mode evaluates each pair of mirror classes in one orientation, chosen per route
    # by timing, so the two routes' counts agree per pair of mirror classes
    assert np.array_equal(cj, cj0) and np.array_equal(_fold_mirrors(ck), _fold_mirrors(ck0))
    dJ, dK = np.abs(J - Jo).max(), np.abs(K - Ko).max()
    _say("  max|dJ| = %.3e  max|dK| = %.3e" % (dJ, dK))
    assert dJ <= 1.0e-7 and dK <= 1.0e-7
    # the FP32 route changes results (it is not silently the FP64 one) ...
    assert np.abs(J - J0).max() > 0.0 and np.abs(K - K0).max() > 0.0
    # ... and with thredp <= thresp it is the FP64 route (equal up to the order of the FP64
    # atomic accumulation)
    J1 = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), 1.0e-12, 1.0e-13).np
    assert np.abs(J1 - J0).max() <= 1.0e-12


def test_mixed_route_ewald_and_nonsymmetric(product_backend, oracle_backend):
    """The FP32 route with erf-attenuated operators and a non-symmetric K call (D24symm False)."""
    ls = product_backend.ls
    _say("mixed route ewald / nonsymmetric")
    mol = FC.molecule("h2o")
    bas = ls.Basis.from_gbs_file(mol, "cc-pvdz")
    rng = np.random.default_rng(5)
    D = rng.standard_normal((bas.nao, bas.nao))
    pl = ls.PairList.build_schwarz(bas, bas, False, 1.0e-14)
    ob = oracle_backend.O.OBasis(bas)
    opl = oracle_backend.O.OPairList(ob, ob, False, 1.0e-14)
    res = ls.ResourceList.build()
    for ew in (FC.COULOMB, FC.EWALD_OPS["MIX"], FC.EWALD_OPS["SR"]):
        J = ls.IntBox.coulomb(res, ls.Ewald(*ew), pl, pl, ls.Tensor.array(D), 1.0e-14, 1.0e-4).np
        K = ls.IntBox.exchange(res, ls.Ewald(*ew), pl, pl, ls.Tensor.array(D), False, 1.0e-14, 1.0e-4).np
        Jo = oracle_backend.O.coulomb(ew, opl, opl, D, 1.0e-14)
        Ko = oracle_backend.O.exchange(ew, opl, opl, D, False, 1.0e-14)
        scale = max(1.0, np.abs(Jo).max())
        assert np.abs(J - Jo).max() <= 1.0e-6 * scale, ew
        assert np.abs(K - Ko).max() <= 1.0e-6 * scale, ew
