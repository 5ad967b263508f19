"""The oracle's per-quartet magnitudes (oracle_quartet_tensor / oracle_quartet_max), the quantity the
learned screen is trained to bound, pinned on the CPU:

* against the oracle's own Coulomb build (itself pinned by the reference's fixtures): on a basis of
  single-primitive shells every basis function is one primitive, and J_ab for a density with one
  non-zero element D_cd = 1 is the integral (ab|cd) itself, for every Cartesian component;
* the (ss|ss) closed form 2 pi^(5/2) K_AB K_CD / (p q sqrt(p+q)) F0(T);
* max over components equals the largest |entry| of the tensor.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

GBS = """cartesian
****
O 0
S 1 1.00
  5.0 1.0
P 1 1.00
  1.2 1.0
D 1 1.00
  0.8 1.0
****
H 0
S 1 1.00
  0.5 1.0
P 1 1.00
  0.7 1.0
****
"""


def _basis(ls, W):
    mol = W.water_cluster(1, 4.0, 5)
    return ls.Basis.from_gbs_str(mol, GBS, name="unc")


def test_quartet_tensor_against_oracle_coulomb(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    bas = _basis(ls, W)
    ob = O.OBasis(bas)
    assert ob.nprim == len(set(int(c) for c in ob.cart))  # one primitive per shell
    opl = O.OPairList(ob, ob, True, 1.0e-30)
    rng = np.random.default_rng(7)
    ew = ([1.0], [-1.0])
    checked = 0
    for _ in range(6):
        P = rng.integers(0, ob.nprim, size=4)
        (T,) = O.quartet_tensor(ew, ob, ob, ob, ob, [P[0]], [P[1]], [P[2]], [P[3]])
        c3, c4 = int(ob.cart[P[2]]), int(ob.cart[P[3]])
        N3, N4 = T.shape[2], T.shape[3]
        c, d = int(rng.integers(0, N3)), int(rng.integers(0, N4))
        D = np.zeros((ob.ncart, ob.ncart))
        D[c3 + c, c4 + d] = 1.0
        J = O.coulomb(ew, opl, opl, D, 1.0e-30)
        c1, c2 = int(ob.cart[P[0]]), int(ob.cart[P[1]])
        ref = J[c1:c1 + T.shape[0], c2:c2 + T.shape[1]]
        assert np.abs(ref).max() > 1.0e-6
        np.testing.assert_allclose(T[:, :, c, d], ref, rtol=1.0e-10, atol=1.0e-13)
        checked += 1
    assert checked == 6


def test_quartet_max_ssss_closed_form(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    bas = _basis(ls, W)
    ob = O.OBasis(bas)
    s = [i for i in range(ob.nprim) if ob.L[i] == 0]
    idx = np.array([(a, b, c, d) for a in s for b in s for c in s for d in s][:60])
    m = O.quartet_max(([1.0], [-1.0]), ob, ob, ob, ob, idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3])
    for (a, b, c, d), v in zip(idx, m):
        ea, eb, ec, ed = ob.e[[a, b, c, d]]
        A, B, C, Dx = ob.xyz[[a, b, c, d]]
        p, q = ea + eb, ec + ed
        Kab = ob.c[a] * ob.c[b] * math.exp(-ea * eb / p * np.sum((A - B) ** 2))
        Kcd = ob.c[c] * ob.c[d] * math.exp(-ec * ed / q * np.sum((C - Dx) ** 2))
        P = (ea * A + eb * B) / p
        Q = (ec * C + ed * Dx) / q
        T = p * q / (p + q) * np.sum((P - Q) ** 2)
        F0 = 1.0 if T < 1e-12 else 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))
        ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0
        assert abs(v - abs(ref)) <= 1.0e-12 * max(1.0, abs(ref))


def test_quartet_max_is_tensor_max(product_backend, oracle_backend):
    ls = product_backend.ls
    O = oracle_backend.O
    from lightspeed_amd import workloads as W
    bas = _basis(ls, W)
    ob = O.OBasis(bas)
    rng = np.random.default_rng(3)
    P = rng.integers(0, ob.nprim, size=(40, 4))
    ew = ([1.0, -0.5], [-1.0, 0.3])
    m = O.quartet_max(ew, ob, ob, ob, ob, *P.T)
    T = O.quartet_tensor(ew, ob, ob, ob, ob, *P.T)
    for v, t in zip(m, T):
        assert v == np.abs(t).max()
