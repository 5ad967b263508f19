"""The autotuner (lsjk_autotune / lsjk_batch_autotune): it changes only how the work is cut into
tasks and units, so J and K after tuning equal J and K before it (1e-12) and the oracle (1e-10);
the tuned table holds one entry per launched J group pair and K class with a candidate value."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRE = 1.0e-12


def test_autotune_single_list(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    mol = W.water_cluster(8, 9.0, 41)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    D = W.random_density(bas.nao, 40, 42)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    J0 = np.array(ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), THRE, THRE).np)
    K0 = np.array(ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, THRE, THRE).np)
    dD = ls.DeviceArray([bas.ncart, bas.ncart])
    dD.upload(D)
    table = ls.autotune(pl, ew, dD.ptr, THRE, THRE)
    kinds = {k for k, _, _, _, _ in table}
    print("tuned entries: %d (kinds %s)" % (len(table), sorted(kinds)), flush=True)
    assert 0 in kinds or 1 in kinds
    assert 2 in kinds and 4 in kinds
    for kind, idx, val, m0, m1 in table:
        assert ((kind < 2 and val in (0.25, 0.5, 1.0, 2.0, 4.0)) or (kind == 2 and val in (3072, 12288, 49152))
                or (kind == 4 and val in (16, 32, 64)))
        assert m1 <= m0 + 1e-9
    for thredp in (THRE, 1e-6):
        J = np.array(ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), THRE, thredp).np)
        K = np.array(ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, THRE, thredp).np)
        tol = 1e-12 if thredp == THRE else 1e-7
        assert np.abs(J - J0).max() < tol and np.abs(K - K0).max() < tol
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, THRE)
    Jo = O.coulomb(([1.0], [-1.0]), opl, opl, D, THRE)
    Ko = O.exchange(([1.0], [-1.0]), opl, opl, D, True, THRE)
    assert np.abs(J0 - Jo).max() < 1e-10 and np.abs(K0 - Ko).max() < 1e-10


def test_autotune_batch(product_backend):
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    mols = []
    for n, seed in ((3, 51), (1, 52), (6, 53), (2, 54)):
        mol = W.water_cluster(n, W.water_cluster_box(n), seed)
        bas = ls.Basis.from_gbs_file(mol, "6-31gs")
        mols.append((bas, W.random_density(bas.nao, 5 * n, seed + 1), ls.PairList.build_schwarz(bas, bas, True, THRE)))
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    Js0, Ks0 = batch.jk(ew, [ls.Tensor.array(D) for _, D, _ in mols], THRE, THRE)
    buf = np.zeros(batch.nelem)
    for off, (bas, D, _) in zip(batch.offsets, mols):
        for i in range(bas.ncart):
            buf[off + i * batch.ld: off + i * batch.ld + bas.ncart] = D[i]
    dD = ls.DeviceArray([batch.nelem])
    dD.upload(buf)
    table = batch.autotune(ew, dD.ptr, THRE, THRE)
    assert len(table) > 0
    Js, Ks = batch.jk(ew, [ls.Tensor.array(D) for _, D, _ in mols], THRE, THRE)
    for a, b, c, d in zip(Js, Js0, Ks, Ks0):
        assert np.abs(np.array(a.np) - np.array(b.np)).max() < 1e-12
        assert np.abs(np.array(c.np) - np.array(d.np)).max() < 1e-12


@pytest.mark.gpu
def test_autotune_mixed_route_kernels(product_backend):
    """lsjk_autotune on a mixed route: per launch, the mixed or the FP64 kernel, whichever timed
    faster (kind 3 entries, value 1 = FP64 kernel).  Results stay within the mixed route's bar of
    the FP64 build, and the K launches moved to the FP64 kernel evaluate no quartet in FP32."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    mol = W.water_cluster(8, 9.0, 43)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    D = W.random_density(bas.nao, 40, 44)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    Dt = ls.Tensor.array(D)
    J0 = np.array(ls.IntBox.coulomb(res, ew, pl, pl, Dt, THRE, THRE).np)
    K0 = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, True, THRE, THRE).np)
    dD = ls.DeviceArray([bas.ncart, bas.ncart])
    dD.upload(D)
    ls.autotune(pl, ew, dD.ptr, THRE, THRE)
    table = ls.autotune(pl, ew, dD.ptr, THRE, 1e-6)
    routes = [(i, v, m0, m1) for k, i, v, m0, m1 in table if k == 3]
    print("mixed-route launches: %d, on the FP64 kernel: %d" % (len(routes), sum(1 for r in routes if r[1] == 1.0)),
          flush=True)
    assert routes
    for i, v, m0, m1 in routes:
        assert v in (0.0, 1.0) and m1 <= m0 + 1e-9
    ls.set_counting(True)
    J = np.array(ls.IntBox.coulomb(res, ew, pl, pl, Dt, THRE, 1e-6).np)
    K = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, True, THRE, 1e-6).np)
    c32 = ls.last_counts_fp32()
    ls.set_counting(False)
    assert np.abs(J - J0).max() <= 1e-7 and np.abs(K - K0).max() <= 1e-7
    for i, v, _, _ in routes:
        if v == 1.0 and 256 <= i < 256 + 625:
            assert c32[i] == 0, i


@pytest.mark.gpu
def test_every_kgroup_size_gives_the_same_exchange(product_backend):
    """The exchange ket-group rows are a launch parameter only: K with 16, 32 and 64 rows per group
    for every ket pair (where the classes fit the LDS limit) equals the default build, on both
    routes and in the symmetric and full modes."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    mol = W.water_cluster(6, 8.0, 47)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    D = W.random_density(bas.nao, 30, 48)
    pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    Dt = ls.Tensor.array(D)
    ref = {(sym, td): np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, sym, THRE, td).np)
           for sym in (True, False) for td in (THRE, 1e-6)}
    for rows in (16, 32, 64):
        applied = 0
        for L in range(3):
            for Lp in range(3):
                try:
                    ls.set_kgroup_rows(pl, L, Lp, rows)
                    applied += 1
                except RuntimeError:
                    ls.set_kgroup_rows(pl, L, Lp, 0)  # does not fit: the default
        print("rows %d applied to %d of 9 ket pairs" % (rows, applied), flush=True)
        assert applied > 0
        for (sym, td), K0 in ref.items():
            K = np.array(ls.IntBox.exchange(res, ew, pl, pl, Dt, sym, THRE, td).np)
            assert np.abs(K - K0).max() < (1e-12 if td == THRE else 1e-9), (rows, sym, td)
    with pytest.raises(RuntimeError):
        ls.set_kgroup_rows(pl, 0, 0, 48)
