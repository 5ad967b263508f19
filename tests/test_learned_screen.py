"""The learned screen's network and its training kernels.

CPU: the numpy oracle's gradient (oracle/mlp_oracle.py) against central finite differences.
GPU: the HIP forward, backward (gradient of the asymmetric squared error) and Adam kernels against
the numpy oracle on the same parameters and data; a short GPU training run lowers the loss.
"""
import numpy as np
import pytest

NI = 13


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, NI))
    y = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * np.abs(X[:, 3])
    return X, y


def test_oracle_gradient_finite_differences(oracle_backend):
    import mlp_oracle as M
    p = M.init(NI, 1) * 0.7
    X, y = _data(40, 2)
    sw = np.random.default_rng(3).random(40) + 0.5
    L, G = M.grad(p, X, y, NI, sw=sw, wu=3.0)
    assert abs(L - M.loss(p, X, y, NI, sw=sw, wu=3.0)) < 1e-12
    rng = np.random.default_rng(4)
    o = M.offsets(NI)
    picks = [o[k][0] + int(rng.integers(0, o[k][1] - o[k][0])) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4")
             for _ in range(3)]
    h = 1e-6
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, X, y, NI, sw=sw, wu=3.0) - M.loss(p - e, X, y, NI, sw=sw, wu=3.0)) / (2 * h)
        assert abs(fd - G[i]) <= 1e-6 * max(1.0, abs(G[i])), (i, fd, G[i])


@pytest.mark.gpu
def test_hip_forward_and_gradient_against_oracle(oracle_backend):
    import mlp_oracle as M
    from lightspeed_amd import screen as S
    p = (M.init(NI, 5) * 0.8).astype(np.float32)
    X, y = _data(333, 6)  # not a multiple of the 64-sample tile
    X = X.astype(np.float32)
    y = y.astype(np.float32)
    sw = (np.random.default_rng(7).random(333) + 0.5).astype(np.float32)
    net = S.MLP(NI, p)
    f = net.forward(X)
    fo, _ = M.forward(p.astype(np.float64), X.astype(np.float64), NI)
    assert np.abs(f - fo).max() <= 1e-5 * max(1.0, np.abs(fo).max())
    idx = np.random.default_rng(8).permutation(333)[:257].astype(np.int32)
    L, G = net.grad(X, y, idx=idx, sw=sw, wu=2.5)
    Lo, Go = M.grad(p.astype(np.float64), X[idx].astype(np.float64), y[idx], NI, sw=sw[idx], wu=2.5)
    assert abs(L - Lo) <= 1e-4 * max(1.0, Lo)
    assert np.abs(G - Go).max() <= 1e-4 * max(1.0, np.abs(Go).max())


@pytest.mark.gpu
def test_hip_adam_and_training(oracle_backend):
    import mlp_oracle as M
    from lightspeed_amd import screen as S
    X, y = _data(4000, 9)
    net = S.MLP.init(NI, 10)
    p0 = net.params().astype(np.float64)
    # one step of the training loop against the numpy Adam step on the same gradient
    L, G = net.grad(X, y, idx=np.arange(1024, dtype=np.int32))
    p1, _, _ = M.adam(p0, G.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1e-3, 0.9, 0.999, 1e-8, 1)
    g = S.DevBuf.of(G, np.float32)
    m = S.DevBuf(net.np, np.float32)
    v = S.DevBuf(net.np, np.float32)
    S.check(S.lib().lsjk_mlp_adam(net.np, net.p.ptr, g.ptr, m.ptr, v.ptr, 1e-3, 0.9, 0.999, 1e-8, 1))
    assert np.abs(net.params() - p1).max() <= 1e-6
    hist = S.MLP.init(NI, 11).train(X, y, epochs=25, batch=512, lr=3e-3)
    assert hist[-1] < 0.25 * hist[0]


# ---------------------------------------------------------------- unit data and the learned route

def _pairs_of_class(ob, opl, La, Lb):
    """All ordered primitive pairs (row, partner) of a symmetric oracle list with L(row) = La,
    L(partner) = Lb, and their bounds."""
    p1, p2, b = [], [], []
    for L1, L2, a, c, bd in opl.classes():
        for i, j, v in zip(a, c, bd):
            for r, q in ((i, j), (j, i)) if i != j else ((i, j),):
                if ob.L[r] == La and ob.L[q] == Lb:
                    p1.append(r)
                    p2.append(q)
                    b.append(v)
    return np.array(p1, np.int32), np.array(p2, np.int32), np.array(b, np.float64)


@pytest.mark.gpu
def test_unit_magnitudes_against_oracle(product_backend, oracle_backend):
    """Per s/p exchange class of (H2O)1/6-31G**: the largest unit value of the GPU data kernel equals
    the oracle's largest max|(ab|cd)| / (bound_ab bound_cd) over every quartet of the class (the units
    of all row pairs together hold every quartet of the class)."""
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    from lightspeed_amd import screen as S
    mol = W.water_cluster(1, 4.0, 21)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, 1.0e-12)
    smin = 1.0e-14
    L = S._bind_screen()
    import ctypes
    for cls in S.SP_CLASSES:
        n = ctypes.c_int()
        S.check(L.lsjk_unit_features(ctypes.c_void_p(pl.device_ptr), *cls, 1e-12, ctypes.byref(n), None))
        mag = np.zeros(n.value, np.float32)
        S.check(L.lsjk_unit_magnitudes(ctypes.c_void_p(pl.device_ptr), *cls, 1e-12, smin, mag.ctypes.data))
        a1, a2, ab = _pairs_of_class(ob, opl, cls[0], cls[1])
        c1, c2, cb = _pairs_of_class(ob, opl, cls[2], cls[3])
        I, J = np.meshgrid(np.arange(len(a1)), np.arange(len(c1)), indexing="ij")
        I, J = I.ravel(), J.ravel()
        keep = ab[I] * cb[J] >= smin
        I, J = I[keep], J[keep]
        m = O.quartet_max(([1.0], [-1.0]), ob, ob, ob, ob, a1[I], a2[I], c1[J], c2[J])
        ref = float(np.max(m / (ab[I] * cb[J])))
        got = float(mag.max())
        print("class %s: %d quartets, max ratio oracle %.6e gpu %.6e" % (cls, len(I), ref, got), flush=True)
        assert abs(got - ref) <= 2e-6 * ref


def _small_screen(S, seed=0):
    X, y, c = S.training_set(sizes=(4, 6, 8), seed0=300)
    return S.train_screen(X, y, c, epochs=12, nbin=4, seed=seed), (X, y, c)


@pytest.mark.gpu
def test_learned_route_small(product_backend, oracle_backend):
    """A screen trained on (H2O)4..8 (6-31G**) routes (H2O)10: K within 1e-7 of the FP64 oracle, fewer
    evaluated quartets than the Schwarz route; a screen that certifies nothing is the plain FP64
    route (same quartets, results to the rounding of the FP64 atomics)."""
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    from lightspeed_amd import screen as S
    scr, (X, y, c) = _small_screen(S)
    print("screen: %d units, certified classes %s" % (len(y), scr.certified()), flush=True)
    assert len(scr.certified()) >= 4
    tau = scr.predict(X, c)
    ok = np.isfinite(scr.delta[c]).any(axis=1)
    # Schwarz stays the judge (tau <= 1); on these small clusters most units hold a near quartet
    # whose ratio to its Schwarz product is about 1, and a part of the far units tighten
    assert np.all(tau <= 1.0) and np.mean(tau[ok] < 0.99) > 0.05
    mol = W.water_cluster(10, 20.0 * (10 / 64.0) ** (1.0 / 3.0), 77)
    bas = ls.Basis.from_gbs_file(mol, "6-31gss")
    D = W.random_density(bas.nao, 50, 78)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, 1.0e-12)
    Ko = O.exchange(([1.0], [-1.0]), opl, opl, D, True, 1.0e-12)
    ew = ls.Ewald([1.0], [-1.0])
    n = bas.ncart
    dD = ls.DeviceArray([n, n])
    dD.upload(D)
    out = {}
    ls.set_counting(True)
    try:
        for name, fn in (("plain", lambda dK: ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, 1e-12, 1e-12, dK.ptr)),
                         ("learned", lambda dK: scr.exchange_device(ew, pl, pl, dD.ptr, True, 1e-12, 1e-12, dK.ptr)),
                         ("learned_mixed", lambda dK: scr.exchange_device(ew, pl, pl, dD.ptr, True, 1e-12, 1e-6,
                                                                          dK.ptr))):
            dK = ls.DeviceArray([n, n])
            fn(dK)
            out[name] = (dK.download(), int(ls.last_counts()[256:].sum()))
    finally:
        ls.set_counting(False)
    for name, (K, cnt) in out.items():
        print("  %s: %d quartets, max|K - K_oracle| %.2e" % (name, cnt, np.abs(K - Ko).max()), flush=True)
    assert np.abs(out["plain"][0] - Ko).max() < 1e-10
    assert np.abs(out["learned"][0] - Ko).max() <= 1e-7
    assert np.abs(out["learned_mixed"][0] - Ko).max() <= 1e-7
    assert out["learned"][1] < out["plain"][1]
    # a screen that certifies no class
    none = S.Screen(scr.params, scr.mean, scr.scale, np.full((625, scr.nbin), np.nan), scr.smin, fedge=scr.fedge)
    dK = ls.DeviceArray([n, n])
    ls.set_counting(True)
    try:
        none.exchange_device(ew, pl, pl, dD.ptr, True, 1e-12, 1e-6, dK.ptr)
        cnt = int(ls.last_counts()[256:].sum())
    finally:
        ls.set_counting(False)
    assert np.abs(dK.download() - out["plain"][0]).max() < 1e-12
    assert cnt == out["plain"][1]


@pytest.mark.gpu
def test_learned_ablation_h2o64(product_backend, oracle_backend):
    """Ablation at (H2O)64/6-31G** (the benchmark's molecule) with the package's trained screen: K
    quartets evaluated and J+K time with and without the screen at the same thresp (FP64 and
    mixed), every route within 1e-7 of the FP64 oracle on a bra-pair sample, and the SCF energy over
    10 iterations from the core guess within 1e-6 Eh of the FP64 route."""
    import os
    import time
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    from lightspeed_amd import screen as S
    from lightspeed_amd import scf
    assert os.path.exists(S.DEFAULT_MODEL), "trained screen missing (python -m lightspeed_amd.screen)"
    scr = S.Screen.load(S.DEFAULT_MODEL)
    mol, bname, nocc = W.config("h2o64_631gss")
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = W.random_density(bas.nao, nocc, 1)
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    ew = ls.Ewald([1.0], [-1.0])
    n = bas.ncart
    dD, dJ, dK = ls.DeviceArray([n, n]), ls.DeviceArray([n, n]), ls.DeviceArray([n, n])
    dD.upload(D)
    routes = {"fp64": (1e-12, None), "mixed": (1e-6, None), "learned": (1e-12, scr), "learned_mixed": (1e-6, scr)}

    def run(r, pl12=pl, symm=True):
        thredp, sc = r
        ls.IntBox.coulomb_device(ew, pl12, pl, dD.ptr, 1e-12, 1e-12 if sc else thredp, dJ.ptr)
        if sc:
            sc.exchange_device(ew, pl12, pl, dD.ptr, symm, 1e-12, thredp, dK.ptr)
        else:
            ls.IntBox.exchange_device(ew, pl12, pl, dD.ptr, symm, 1e-12, thredp, dK.ptr)

    stats = {}
    for name, r in routes.items():
        ls.set_counting(True)
        dJ.zero()
        dK.zero()
        run(r)
        nk = int(ls.last_counts()[256:].sum())
        ls.set_counting(False)
        run(r)  # warm
        ls.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            run(r)
        ls.synchronize()
        stats[name] = (nk, (time.perf_counter() - t0) / 3 * 1e3)
    for name, (nk, ms) in stats.items():
        print("  %-14s K quartets %d  J+K %.2f ms" % (name, nk, ms), flush=True)
    assert stats["learned"][0] <= stats["fp64"][0]
    # against the FP64 oracle on every 64th bra pair of each class (whole ket list, D24symm off)
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, 1.0e-12)
    sub, osub = pl.subsample(64), opl.subsample(64)
    Jo = O.coulomb(([1.0], [-1.0]), osub, opl, D, 1e-12)
    Ko = O.exchange(([1.0], [-1.0]), osub, opl, D, False, 1e-12)
    for name, r in routes.items():
        dJ.zero()
        dK.zero()
        run(r, pl12=sub, symm=False)
        eJ, eK = np.abs(dJ.download() - Jo).max(), np.abs(dK.download() - Ko).max()
        print("  %-14s vs oracle (sample): max|dJ| %.2e max|dK| %.2e" % (name, eJ, eK), flush=True)
        assert eJ <= 1e-7 and eK <= 1e-7
    # SCF energy over 10 iterations from the core guess
    ints = scf.Integrals(bas, 1.0e-12)
    e64 = scf.RHF(mol, ints, thresp=1e-12, thredp=1e-12, fixed_iterations=10)
    e64.run()
    sints = scf.ScreenedIntegrals(bas, scr, 1.0e-12)
    eln = scf.RHF(mol, sints, thresp=1e-12, thredp=1e-6, fixed_iterations=10)
    eln.run()
    dE = max(abs(a[1] - b[1]) for a, b in zip(e64.history, eln.history))
    print("  SCF 10 iterations: max |dE| learned+mixed vs fp64 = %.3e" % dE, flush=True)
    assert dE <= 1e-6
