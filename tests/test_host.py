"""Host-side logic and the drop-in boundary (CPU only; no GPU calls are made).

  * liblsjk.so loads and exports every function include/lsjk.h declares;
  * the host PairList::build_schwarz agrees bit for bit with the oracle's restatement
    (indices, float bounds, class order) on symmetric, non-symmetric and mixed-basis lists;
  * loaders: primitive normalisation (unit self-overlap of every contracted shell), basis sizes of
    the benchmark workloads, water-cluster geometry constraints;
  * the reference's error behaviour: shape errors, D24symm with unequal bases, no GPU.
"""
import ctypes
import math
import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import lightspeed_amd as ls  # noqa: E402
from lightspeed_amd import loaders, workloads  # noqa: E402
import oracle as O  # noqa: E402
import fixture_cases as FC  # noqa: E402


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "lsjk.h")).read()
    names = re.findall(r"^\s*(?:int|const char\*)\s+(lsjk_\w+)\s*\(", hdr, re.M)
    assert len(names) >= 20
    lib = ctypes.CDLL(ls.NATIVE_LIBRARY)
    for n in names:
        assert hasattr(lib, n), n


def test_device_count_without_gpu_is_zero_or_more():
    lib = ctypes.CDLL(ls.NATIVE_LIBRARY)
    c = ctypes.c_int(-1)
    assert lib.lsjk_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0


def _compare_pairlists(pl, opl):
    cls = opl.classes()
    assert len(pl.pairlists) == len(cls)
    for l, (L1, L2, p1, p2, b) in zip(pl.pairlists, cls):
        a1, a2, ab = l.arrays()
        assert (l.L1, l.L2) == (L1, L2)
        assert np.array_equal(a1, p1) and np.array_equal(a2, p2) and np.array_equal(ab, b)


@pytest.mark.parametrize("basname", ["sto-3g", "6-31gs", "cc-pvdz", "6-31gss"])
@pytest.mark.parametrize("symmetric", [True, False])
def test_pairlist_matches_oracle(basname, symmetric):
    mol = workloads.water_cluster(6, 8.0, 11)
    bas = ls.Basis.from_gbs_file(mol, basname)
    for thre in (1e-14, 1e-10):
        pl = ls.PairList.build_schwarz(bas, bas, symmetric, thre)
        ob = O.OBasis(bas)
        _compare_pairlists(pl, O.OPairList(ob, ob, symmetric, thre))


def test_pairlist_mixed_bases_matches_oracle():
    m1, m2 = FC.molecule("h2o"), FC.molecule("nh3")
    b1 = ls.Basis.from_gbs_file(m1, "cc-pvdz")
    b2 = ls.Basis.from_gbs_file(m2, "sto-3g")
    pl = ls.PairList.build_schwarz(b1, b2, False, 1e-14)
    _compare_pairlists(pl, O.OPairList(O.OBasis(b1), O.OBasis(b2), False, 1e-14))


@pytest.mark.parametrize("basname", ["sto-3g", "6-31g", "6-31gs", "6-31gss", "cc-pvdz", "3-21g"])
def test_contracted_shells_are_normalised(basname):
    """sum_kl c_k c_l (pi/(e_k+e_l))^{3/2} (2L-1)!!/(2(e_k+e_l))^L = 1 for the xx..x component."""
    mol = workloads.water_cluster(1, 5.0, 1)
    bas = ls.Basis.from_gbs_file(mol, basname)
    shells = {}
    for p in bas.primitives:
        shells.setdefault(p.shellIdx, []).append(p)
    for prims in shells.values():
        L = prims[0].L
        df = 1
        for l in range(1, L + 1):
            df *= 2 * l - 1
        s = 0.0
        for a in prims:
            for b in prims:
                p = a.e + b.e
                s += a.c * b.c * (math.pi / p) ** 1.5 * df / (2 * p) ** L
        assert abs(s - 1.0) < 1e-10, (basname, L)


def test_workload_sizes():
    sizes = {"h2o10_631g": (130, 130, 220), "h2o64_631gss": (1600, 1600, 1600),
             "c20h42_ccpvdz": (490, 510, None)}
    for name, (nao, ncart, nprim) in sizes.items():
        mol, bname, nocc = workloads.config(name)
        bas = ls.Basis.from_gbs_file(mol, bname)
        assert bas.nao == nao and bas.ncart == ncart
        if nprim is not None:
            assert bas.nprim == nprim


def test_water_cluster_geometry():
    mol = workloads.water_cluster(20, 15.0, 5)
    xyz = mol.xyz * loaders.ANG_PER_BOHR
    O_ = xyz[0::3]
    for i in range(20):
        o, h1, h2 = xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]
        assert abs(np.linalg.norm(h1 - o) - 0.96) < 1e-12
        assert abs(np.linalg.norm(h2 - o) - 0.96) < 1e-12
        c = np.dot(h1 - o, h2 - o) / 0.96 ** 2
        assert abs(math.degrees(math.acos(c)) - 104.5) < 1e-9
    d = np.linalg.norm(O_[:, None, :] - O_[None, :, :], axis=-1) + np.eye(20) * 99
    assert d.min() >= 2.8 and np.all(O_ >= 0) and np.all(O_ <= 15.0)


def test_density_is_orthogonal_projector():
    D = workloads.random_density(40, 9, 3)
    assert np.allclose(D, D.T) and np.allclose(D @ D, D, atol=1e-12) and abs(np.trace(D) - 9) < 1e-12


def test_tensor_roundtrip_and_axpby():
    a = np.arange(12.0).reshape(3, 4)
    t = ls.Tensor.array(a)
    assert t.shape == [3, 4] and np.array_equal(t.np, a)
    u = ls.Tensor.array(np.ones((3, 4)))
    t.axpby(u, 2.0, 0.5)
    assert np.allclose(t.np, 0.5 * a + 2.0)
    with pytest.raises(RuntimeError):
        t.shape_error([4, 3])


def test_pure_transform_roundtrip_consistency():
    """pureToCart2 then the adjoint cartToPure2 equals C C^T D C C^T per shell (oracle matrices)."""
    mol = FC.molecule("h2o")
    bas = ls.Basis.from_gbs_file(mol, "cc-pvdz")
    rng = np.random.default_rng(0)
    D = rng.standard_normal((bas.nao, bas.nao))
    T = O.OBasis(bas).transform()
    Dc = np.array(ls.PureTransform.pureToCart2(bas, bas, ls.Tensor.array(D)).np)
    assert np.allclose(Dc, T.T @ D @ T, atol=1e-13)
    Jp = np.array(ls.PureTransform.cartToPure2(bas, bas, ls.Tensor.array(Dc)).np)
    assert np.allclose(Jp, T @ Dc @ T.T, atol=1e-13)


def test_solid_harmonics_closed_forms():
    C2 = ls.PureTransform.solid_harmonics(2)  # rows 0, 1c, 1s, 2c, 2s; columns xx xy xz yy yz zz
    s3 = math.sqrt(3.0)
    assert np.allclose(C2[0], [-0.5, 0, 0, -0.5, 0, 1.0])
    assert np.allclose(C2[1], [0, 0, s3, 0, 0, 0])
    assert np.allclose(C2[2], [0, 0, 0, 0, s3, 0])
    assert np.allclose(C2[3], [0.5 * s3, 0, 0, -0.5 * s3, 0, 0])
    assert np.allclose(C2[4], [0, s3, 0, 0, 0, 0])
    C1 = ls.PureTransform.solid_harmonics(1)  # z, x, y
    assert np.allclose(C1, [[0, 0, 1], [1, 0, 0], [0, 1, 0]])
    for L in range(4):
        assert np.allclose(ls.PureTransform.solid_harmonics(L), O.solid_harmonics(L), atol=1e-15)


def test_intbox_errors_without_gpu():
    mol = FC.molecule("h2o")
    bas = ls.Basis.from_gbs_file(mol, "6-31gs")
    pl = ls.PairList.build_schwarz(bas, bas, True, 1e-14)
    res = ls.ResourceList.build_cpu()
    ew = ls.Ewald([1.0], [-1.0])
    D = ls.Tensor.array(np.eye(bas.nao))
    with pytest.raises(RuntimeError, match="HIP backend unavailable"):
        ls.IntBox.coulomb(res, ew, pl, pl, D, 1e-14, 1e-14)
    with pytest.raises(RuntimeError, match="shape"):
        ls.IntBox.coulomb(res, ew, pl, pl, ls.Tensor.array(np.eye(3)), 1e-14, 1e-14)
    other = ls.Basis.from_gbs_file(FC.molecule("nh3"), "sto-3g")
    pl2 = ls.PairList.build_schwarz(bas, other, False, 1e-14)
    with pytest.raises(RuntimeError, match="D24symm"):
        ls.IntBox.exchange(res, ew, pl2, pl, ls.Tensor.array(np.zeros((other.nao, bas.nao))), True, 1e-14, 1e-14)


def test_basis_dense_index_validation():
    P = ls.Primitive
    with pytest.raises(RuntimeError, match="primIdx"):
        ls.Basis("b", [P(1.0, 1.0, 0, 0, 0, 1.0, 0, False, 0, 0, 1, 0, 0)])
    with pytest.raises(RuntimeError, match="aoIdx"):
        ls.Basis("b", [P(1.0, 1.0, 0, 0, 0, 1.0, 0, False, 0, 0, 0, 0, 0),
                       P(1.0, 1.0, 0, 0, 0, 1.0, 0, False, 2, 1, 1, 1, 0)])


def test_pairlist_symmetric_requires_equal_bases():
    b1 = ls.Basis.from_gbs_file(FC.molecule("h2o"), "sto-3g")
    b2 = ls.Basis.from_gbs_file(FC.molecule("nh3"), "sto-3g")
    with pytest.raises(RuntimeError, match="not symmetric"):
        ls.PairList.build_schwarz(b1, b2, True, 1e-14)


def test_bench_refuses_several_gpus():
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_pure_transforms_per_block_against_dense(product_backend, oracle_backend):
    """PureTransform.pureToCart2 / cartToPure2 work shell pair by shell pair; they equal the dense
    products with the oracle's nao x ncart transform matrix (C4H10 / cc-pVDZ, spherical d)."""
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    bas = ls.Basis.from_gbs_file(W.alkane(4), "cc-pvdz")
    assert bas.has_pure and bas.nao < bas.ncart
    T = O.OBasis(bas).transform()
    rng = np.random.default_rng(0)
    P = rng.standard_normal((bas.nao, bas.nao))
    C = np.array(ls.PureTransform.pureToCart2(bas, bas, ls.Tensor.array(P)).np)
    assert np.abs(C - T.T @ P @ T).max() < 1e-13
    Kc = rng.standard_normal((bas.ncart, bas.ncart))
    Q0 = rng.standard_normal((bas.nao, bas.nao))
    Q = np.array(ls.PureTransform.cartToPure2(bas, bas, ls.Tensor.array(Kc), ls.Tensor.array(Q0.copy())).np)
    assert np.abs(Q - (Q0 + T @ Kc @ T.T)).max() < 1e-13
