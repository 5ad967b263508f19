"""lsjk_coulomb_noalloc (the J build whose per-call buffers live in the library's fixed device
scratch array) against lsjk_coulomb on the same inputs, through the C-ABI (ctypes), on two calls
in a row with the second on a smaller input: (H2O)6 then (H2O)2, 6-31G**, Schwarz 1e-12, random
orthonormal C.  Both the FP64 route and the mixed route (FP32 copy of the ket densities, also in
the scratch) are compared; the two entry points share their kernels, so they agree to the
rounding of the FP64 atomics (1e-12 absolute).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRE = 1.0e-12


def _case(lib, ls, W, nwat, seed):
    mol = W.water_cluster(nwat, 7.0, seed)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    D = W.random_density(bas.nao, 5 * nwat, seed)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    hb = lib.basis(bas)
    hpl = lib.pairlist(hb, hb, pl)
    return bas, D, hb, hpl


def test_noalloc_coulomb_matches_coulomb(product_backend):
    import capi_ctypes as C
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    lib = C.Lib(0)
    ew = C.Lib.ewald([1.0], [-1.0])
    for nwat, seed in ((6, 11), (2, 12)):
        bas, D, hb, hpl = _case(lib, ls, W, nwat, seed)
        n = bas.ncart
        for thredp in (THRE, 1.0e-6):
            print("noalloc J (H2O)%d/6-31G** n=%d thredp=%g" % (nwat, n, thredp), flush=True)
            sys.stdout.flush()
            dD = lib.upload(D)
            dJ = lib.zeros((n, n))
            dJn = lib.zeros((n, n))
            try:
                lib.coulomb(hpl, hpl, dD, ew, THRE, thredp, dJ)
                lib.coulomb_noalloc(hpl, hpl, dD, ew, THRE, thredp, dJn)
                J = lib.download(dJ, (n, n))
                Jn = lib.download(dJn, (n, n))
            finally:
                for d in (dD, dJ, dJn):
                    lib.free(d)
            assert np.abs(J).max() > 0.1
            assert np.abs(J - Jn).max() < 1.0e-12
        lib.check(lib.L.lsjk_pairlist_destroy(hpl))
        lib.check(lib.L.lsjk_basis_destroy(hb))
