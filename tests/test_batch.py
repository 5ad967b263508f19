"""Batched J + K builds of many independent molecules (JKBatch, lsjk_batch_*): one launch per class
over all molecules, replayed from a captured hipGraph after the first call.

* Small batches (water clusters of 1..6 waters, 6-31G*, and mixed bases) against each molecule's
  own IntBox.coulomb / IntBox.exchange (same library, 1e-12) and the CPU FP64 oracle (1e-10).
* Graph replays: the same device buffers zeroed and rebuilt several times, with eager library
  work and synchronisations in between; every replay equals the first (eager) call's result.
* The mixed route of a batch: within 1e-7 of the FP64 route, per molecule.
* The batched workload itself: 256 seeded (H2O)_n, n uniform in 1..16, 6-31G* (nocc = 5n), all
  J and K from one graph replay, each molecule equal to its single-molecule build at 1e-12.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRE = 1.0e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _molecules(ls, W, specs, basis="6-31gs"):
    out = []
    for n, seed in specs:
        mol = W.water_cluster(n, W.water_cluster_box(n), seed)
        bas = ls.Basis.from_gbs_file(mol, basis)
        D = W.random_density(bas.nao, 5 * n, seed + 1)
        pl = ls.PairList.build_schwarz(bas, bas, True, THRE)
        out.append((bas, D, pl))
    return out


def _single(ls, res, ew, pl, D, thredp=THRE):
    J = ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(D), THRE, thredp).np
    K = ls.IntBox.exchange(res, ew, pl, pl, ls.Tensor.array(D), True, THRE, thredp).np
    return np.array(J), np.array(K)


def test_batch_small_against_single_and_oracle(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    specs = [(1, 11), (3, 12), (2, 13), (6, 14), (1, 15), (4, 16), (5, 17), (2, 18)]
    mols = _molecules(ls, W, specs)
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    _say("batch of %d molecules (nao %s)" % (len(mols), [b.nao for b, _, _ in mols]))
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    assert batch.nmol == len(mols) and batch.ld == max(b.ncart for b, _, _ in mols)
    Js, Ks = batch.jk(ew, [ls.Tensor.array(D) for _, D, _ in mols], THRE, THRE)
    for (bas, D, pl), J, K in zip(mols, Js, Ks):
        J, K = np.array(J.np), np.array(K.np)
        Js1, Ks1 = _single(ls, res, ew, pl, D)
        ob = O.OBasis(bas)
        opl = O.OPairList(ob, ob, True, THRE)
        Jo = O.coulomb(([1.0], [-1.0]), opl, opl, D, THRE)
        Ko = O.exchange(([1.0], [-1.0]), opl, opl, D, True, THRE)
        d1 = max(np.abs(J - Js1).max(), np.abs(K - Ks1).max())
        do = max(np.abs(J - Jo).max(), np.abs(K - Ko).max())
        _say("  nao %3d: |batch - single| %.1e  |batch - oracle| %.1e" % (bas.nao, d1, do))
        assert np.abs(Jo).max() > 0.1 and np.abs(Ko).max() > 0.1
        assert d1 < 1.0e-12 and do < 1.0e-10


def test_batch_mixed_bases_and_operator(product_backend, oracle_backend):
    """Molecules with different bases (6-31G**, cc-pVDZ spherical) and an erf-attenuated operator."""
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    mols = _molecules(ls, W, [(2, 21)], "6-31gss") + _molecules(ls, W, [(1, 22), (2, 23)], "cc-pvdz")
    res = ls.ResourceList.build()
    scales, omegas = [1.0, -0.6], [-1.0, 0.4]
    ew = ls.Ewald(scales, omegas)
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    Js, Ks = batch.jk(ew, [ls.Tensor.array(D) for _, D, _ in mols], THRE, THRE)
    for (bas, D, pl), J, K in zip(mols, Js, Ks):
        J, K = np.array(J.np), np.array(K.np)
        Js1, Ks1 = _single(ls, res, ew, pl, D)
        ob = O.OBasis(bas)
        opl = O.OPairList(ob, ob, True, THRE)
        Jo = O.coulomb((scales, omegas), opl, opl, D, THRE)
        Ko = O.exchange((scales, omegas), opl, opl, D, True, THRE)
        d1 = max(np.abs(J - Js1).max(), np.abs(K - Ks1).max())
        do = max(np.abs(J - Jo).max(), np.abs(K - Ko).max())
        _say("  %s nao %3d: |batch - single| %.1e  |batch - oracle| %.1e" % (bas.name, bas.nao, d1, do))
        assert d1 < 1.0e-12 and do < 1.0e-10


def _pack(batch, mats):
    buf = np.zeros(batch.nelem)
    for off, M in zip(batch.offsets, mats):
        n = M.shape[0]
        for i in range(n):
            buf[off + i * batch.ld: off + i * batch.ld + n] = M[i]
    return buf


def _unpack(batch, buf, sizes):
    out = []
    for off, n in zip(batch.offsets, sizes):
        out.append(np.array([buf[off + i * batch.ld: off + i * batch.ld + n] for i in range(n)]))
    return out


def test_batch_graph_replays(product_backend):
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    mols = _molecules(ls, W, [(3, 31), (1, 32), (5, 33), (2, 34)])
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    sizes = [b.ncart for b, _, _ in mols]
    dD = ls.DeviceArray([batch.nelem])
    dJ = ls.DeviceArray([batch.nelem])
    dK = ls.DeviceArray([batch.nelem])
    dD.upload(_pack(batch, [D for _, D, _ in mols]))
    results = []
    for rep in range(4):
        for thredp in (THRE, 1.0e-6):
            dJ.zero()
            dK.zero()
            batch.jk_device(ew, dD.ptr, THRE, thredp, dJ.ptr, dK.ptr)
            results.append((thredp, dJ.download(), dK.download()))
        # eager library work between the replays
        _single(ls, res, ew, mols[0][2], mols[0][1])
        ls.synchronize()
    ref = {thredp: (J, K) for thredp, J, K in results[:2]}
    for i, (thredp, J, K) in enumerate(results):
        J0, K0 = ref[thredp]
        d = max(np.abs(J - J0).max(), np.abs(K - K0).max())
        _say("  call %d thredp %g: |replay - first| %.1e" % (i, thredp, d))
        assert d < 1.0e-12
    # the replayed results are the molecules' own
    for (bas, D, pl), J, K in zip(mols, _unpack(batch, results[-2][1], sizes), _unpack(batch, results[-2][2], sizes)):
        J1, K1 = _single(ls, res, ew, pl, D)
        assert max(np.abs(J - J1).max(), np.abs(K - K1).max()) < 1.0e-12
    # mixed route of the batch within 1e-7 of FP64, per molecule
    Jm = _unpack(batch, results[-1][1], sizes)
    J6 = _unpack(batch, results[-2][1], sizes)
    Km = _unpack(batch, results[-1][2], sizes)
    K6 = _unpack(batch, results[-2][2], sizes)
    for a, b, c, d in zip(Jm, J6, Km, K6):
        assert np.abs(a - b).max() <= 1.0e-7 and np.abs(c - d).max() <= 1.0e-7


def test_batched_workload_256(product_backend):
    """256 seeded (H2O)_n, n uniform in 1..16, 6-31G*: one graph replay against the single builds."""
    ls = product_backend.ls
    from lightspeed_amd import workloads as W
    res = ls.ResourceList.build()
    ew = ls.Ewald([1.0], [-1.0])
    work = W.water_batch(256)
    mols = []
    for mol, nocc, dseed in work:
        bas = ls.Basis.from_gbs_file(mol, "6-31gs")
        D = W.random_density(bas.nao, nocc, dseed)
        mols.append((bas, D, ls.PairList.build_schwarz(bas, bas, True, THRE)))
    sizes = [b.ncart for b, _, _ in mols]
    _say("256 molecules: %d functions in all, largest %d" % (sum(sizes), max(sizes)))
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    dD = ls.DeviceArray([batch.nelem])
    dJ = ls.DeviceArray([batch.nelem])
    dK = ls.DeviceArray([batch.nelem])
    dD.upload(_pack(batch, [D for _, D, _ in mols]))
    batch.jk_device(ew, dD.ptr, THRE, THRE, dJ.ptr, dK.ptr)  # eager + capture
    dJ.zero()
    dK.zero()
    batch.jk_device(ew, dD.ptr, THRE, THRE, dJ.ptr, dK.ptr)  # replay
    ms = ls.last_timing()[0]
    _say("  graph replay %.1f ms (%.0f molecules/s)" % (ms, 256.0 / (ms * 1e-3)))
    Js = _unpack(batch, dJ.download(), sizes)
    Ks = _unpack(batch, dK.download(), sizes)
    worst = 0.0
    for (bas, D, pl), J, K in zip(mols, Js, Ks):
        J1, K1 = _single(ls, res, ew, pl, D)
        assert np.abs(J1).max() > 0.1
        worst = max(worst, np.abs(J - J1).max(), np.abs(K - K1).max())
    _say("  max |batch - single| over the 256 molecules: %.1e" % worst)
    assert worst < 1.0e-12
