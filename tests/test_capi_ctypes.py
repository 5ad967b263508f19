"""The C-ABI driven directly through ctypes (no C++ or pybind layer), as a caller without the
host library binds it: lsjk_basis_create / lsjk_pairlist_create from plain arrays, then
lsjk_coulomb and lsjk_exchange (D24symm) on device buffers, against the CPU FP64 oracle at 1e-10
and against IntBox.coulomb / IntBox.exchange (same library) at 1e-12.  (H2O)3 / 6-31G**, Schwarz
1e-12, random orthonormal C (nocc 15).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRE = 1.0e-12


def test_ctypes_coulomb_exchange_against_oracle(product_backend, oracle_backend):
    import capi_ctypes as C
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    print("ctypes C-ABI (H2O)3/6-31G**", flush=True)
    sys.stdout.flush()
    mol = W.water_cluster(3, 6.0, 3)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    assert bas.nao == bas.ncart
    D = W.random_density(bas.nao, 15, 3)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    lib = C.Lib(0)
    hb = lib.basis(bas)
    hpl = lib.pairlist(hb, hb, pl)
    n = bas.ncart
    ew = C.Lib.ewald([1.0], [-1.0])
    dD = lib.upload(D)
    dJ = lib.zeros((n, n))
    dK = lib.zeros((n, n))
    try:
        lib.coulomb(hpl, hpl, dD, ew, THRE, THRE, dJ)
        lib.exchange(hpl, hpl, dD, True, ew, THRE, THRE, dK)
        J = lib.download(dJ, (n, n))
        K = lib.download(dK, (n, n))
    finally:
        for d in (dD, dJ, dK):
            lib.free(d)
        lib.check(lib.L.lsjk_pairlist_destroy(hpl))
        lib.check(lib.L.lsjk_basis_destroy(hb))
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, THRE)
    Jo = O.coulomb(([1.0], [-1.0]), opl, opl, D, THRE)
    Ko = O.exchange(([1.0], [-1.0]), opl, opl, D, True, THRE)
    assert np.abs(Jo).max() > 0.1 and np.abs(Ko).max() > 0.1
    assert np.abs(J - Jo).max() < 1.0e-10
    assert np.abs(K - Ko).max() < 1.0e-10
    res = ls.ResourceList.build()
    ewl = ls.Ewald([1.0], [-1.0])
    Jh = ls.IntBox.coulomb(res, ewl, pl, pl, ls.Tensor.array(D), THRE, THRE).np
    Kh = ls.IntBox.exchange(res, ewl, pl, pl, ls.Tensor.array(D), True, THRE, THRE).np
    assert np.abs(J - Jh).max() < 1.0e-12
    assert np.abs(K - Kh).max() < 1.0e-12
