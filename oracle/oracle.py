"""TEST INFRASTRUCTURE ONLY -- Python front end of the CPU FP64 oracle (oracle/jk_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker.  It mirrors the reference's IntBox::coulomb / IntBox::exchange CPU path
(src/cpp/intbox/cpu/coulomb.cpp:14-196, src/cpp/intbox/cpu/exchange.cpp:17-355) including the
pure <-> Cartesian transforms around the Cartesian kernels
(src/cpp/core/pure_transform.cpp:9-247) and the real solid harmonics of
src/cpp/core/am.cpp:74-203 (Helgaker-Jorgensen-Olsen eqs. 6.4.70-6.4.73, re-derived here).

Inputs are any basis objects exposing `.primitives` records with fields
c, e, x, y, z, c0, L, is_pure, aoIdx, cartIdx, primIdx, shellIdx, atomIdx.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_jk.so")
_lib = None

MAXCLS = 64


class _CBasis(ctypes.Structure):
    _fields_ = [("nprim", ctypes.c_int), ("ncart", ctypes.c_int),
                ("c", ctypes.c_void_p), ("e", ctypes.c_void_p), ("xyz", ctypes.c_void_p),
                ("L", ctypes.c_void_p), ("cart", ctypes.c_void_p)]


class _CPairList(ctypes.Structure):
    _fields_ = [("symmetric", ctypes.c_int), ("npair", ctypes.c_int),
                ("p1", ctypes.POINTER(ctypes.c_int)), ("p2", ctypes.POINTER(ctypes.c_int)),
                ("bound", ctypes.POINTER(ctypes.c_float)), ("ncls", ctypes.c_int),
                ("cls_L1", ctypes.c_int * MAXCLS), ("cls_L2", ctypes.c_int * MAXCLS),
                ("cls_start", ctypes.c_int * (MAXCLS + 1))]


def build(force=False):
    """Compile liboracle_jk.so with gcc (OpenMP)."""
    src = os.path.join(_HERE, "jk_oracle.c")
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_jk.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB)
        L.oracle_pairlist_build.argtypes = [ctypes.POINTER(_CBasis), ctypes.POINTER(_CBasis), ctypes.c_int,
                                            ctypes.c_float, ctypes.POINTER(_CPairList)]
        L.oracle_pairlist_free.argtypes = [ctypes.POINTER(_CPairList)]
        L.oracle_coulomb.argtypes = [ctypes.POINTER(_CBasis)] * 4 + [
            ctypes.POINTER(_CPairList), ctypes.POINTER(_CPairList), ctypes.c_void_p, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_exchange.argtypes = [ctypes.POINTER(_CBasis)] * 4 + [
            ctypes.POINTER(_CPairList), ctypes.POINTER(_CPairList), ctypes.c_int, ctypes.c_void_p,
            ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
            ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_rys.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_overlap_kinetic.argtypes = [ctypes.POINTER(_CBasis)] * 2 + [ctypes.POINTER(_CPairList),
                                                                            ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_potential.argtypes = [ctypes.POINTER(_CBasis)] * 2 + [
            ctypes.POINTER(_CPairList), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_void_p]
        for f in (L.oracle_quartet_max, L.oracle_quartet_tensor):
            f.argtypes = [ctypes.POINTER(_CBasis)] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 4 + [
                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


# ------------------------------------------------------------------ basis data

class OBasis:
    """Flat arrays of a basis (the oracle's own copy of the primitive data)."""

    def __init__(self, basis):
        prims = list(basis.primitives)
        self.src = basis
        self.nprim = len(prims)
        self.c = np.array([p.c for p in prims], dtype=np.float64)
        self.e = np.array([p.e for p in prims], dtype=np.float64)
        self.xyz = np.array([[p.x, p.y, p.z] for p in prims], dtype=np.float64).reshape(-1, 3)
        self.L = np.array([p.L for p in prims], dtype=np.int32)
        self.cart = np.array([p.cartIdx for p in prims], dtype=np.int32)
        self.ao = np.array([p.aoIdx for p in prims], dtype=np.int32)
        self.shell = np.array([p.shellIdx for p in prims], dtype=np.int32)
        self.pure = np.array([bool(p.is_pure) for p in prims], dtype=bool)
        self.key = tuple((p.c, p.e, p.x, p.y, p.z, p.c0, p.L, bool(p.is_pure), p.aoIdx, p.cartIdx,
                          p.primIdx, p.shellIdx, p.atomIdx) for p in prims)
        if self.nprim:
            last = prims[-1]
            self.ncart = last.cartIdx + (last.L + 1) * (last.L + 2) // 2
            self.nao = last.aoIdx + ((2 * last.L + 1) if last.is_pure else (last.L + 1) * (last.L + 2) // 2)
        else:
            self.ncart = self.nao = 0
        self.cstruct = _CBasis(self.nprim, self.ncart, self.c.ctypes.data, self.e.ctypes.data,
                               self.xyz.ctypes.data, self.L.ctypes.data, self.cart.ctypes.data)

    def equivalent(self, other):
        """Basis::equivalent, src/lightspeed/basis.hpp:583-592."""
        return self.nao == other.nao and self.ncart == other.ncart and self.key == other.key

    def transform(self):
        """nao x ncart matrix T with phi_pure = T phi_cart (identity blocks for Cartesian shells)."""
        T = np.zeros((self.nao, self.ncart))
        seen = set()
        for i in range(self.nprim):
            s = self.shell[i]
            if s in seen:
                continue
            seen.add(s)
            L, a, c = int(self.L[i]), int(self.ao[i]), int(self.cart[i])
            nc = (L + 1) * (L + 2) // 2
            if self.pure[i]:
                T[a:a + 2 * L + 1, c:c + nc] = solid_harmonics(L)
            else:
                T[a:a + nc, c:c + nc] = np.eye(nc)
        return T


def _cart_index(L):
    idx = {}
    k = 0
    for i in range(L + 1):
        l = L - i
        for j in range(i + 1):
            idx[(l, i - j, j)] = k
            k += 1
    return idx


def solid_harmonics(L):
    """(2L+1) x ncart(L) real solid harmonic coefficients, order (0, 1c, 1s, 2c, 2s, ...),
    from the recursions of Helgaker-Jorgensen-Olsen 6.4.70-6.4.73 (as am.cpp:74-203 does)."""
    def add(a, b, f=1.0):
        out = dict(a)
        for k, v in b.items():
            out[k] = out.get(k, 0.0) + f * v
        return out

    def mul(poly, dx, dy, dz, f=1.0):
        return {(k[0] + dx, k[1] + dy, k[2] + dz): f * v for k, v in poly.items()}

    Sc = [{(0, 0, 0): 1.0}]
    Ss = [{}]
    for l in range(L):
        Nc = math.sqrt((2.0 if l == 0 else 1.0) * (2.0 * l + 1) / (2.0 * l + 2.0))
        Ns = 0.0 if l == 0 else Nc
        Sc.append(add(mul(Sc[l], 1, 0, 0, Nc), mul(Ss[l], 0, 1, 0, -Ns)))
        Ss.append(add(mul(Sc[l], 0, 1, 0, Nc), mul(Ss[l], 1, 0, 0, Ns)))
    rows = [None] * (2 * L + 1)
    for m in range(L + 1):
        cur_c, cur_s = Sc[m], Ss[m]
        prev_c, prev_s = {}, {}
        for l in range(m, L):
            a = (2.0 * l + 1.0) / math.sqrt((l + m + 1.0) * (l - m + 1.0))
            b = math.sqrt(((l + m) * float(l - m)) / ((l + m + 1.0) * (l - m + 1.0)))
            r2c = add(add(mul(prev_c, 2, 0, 0), mul(prev_c, 0, 2, 0)), mul(prev_c, 0, 0, 2))
            r2s = add(add(mul(prev_s, 2, 0, 0), mul(prev_s, 0, 2, 0)), mul(prev_s, 0, 0, 2))
            nc = add(mul(cur_c, 0, 0, 1, a), r2c, -b)
            ns = add(mul(cur_s, 0, 0, 1, a), r2s, -b)
            prev_c, prev_s, cur_c, cur_s = cur_c, cur_s, nc, ns
        if m == 0:
            rows[0] = cur_c
        else:
            rows[2 * m - 1] = cur_c
            rows[2 * m] = cur_s
    idx = _cart_index(L)
    C = np.zeros((2 * L + 1, len(idx)))
    for r, poly in enumerate(rows):
        for k, v in poly.items():
            if abs(v) > 1e-15:
                C[r, idx[k]] += v
    return C


# ------------------------------------------------------------------ pair lists

class OPairList:
    """A Schwarz pair list built by the oracle (PairList::build_schwarz, pair_list.cpp:29-177)."""

    def __init__(self, basis1, basis2, symmetric, threpq):
        self.b1 = basis1 if isinstance(basis1, OBasis) else OBasis(basis1)
        self.b2 = basis2 if isinstance(basis2, OBasis) else OBasis(basis2)
        self.symmetric = bool(symmetric)
        self.threpq = threpq
        self.c = _CPairList()
        rc = lib().oracle_pairlist_build(ctypes.byref(self.b1.cstruct), ctypes.byref(self.b2.cstruct),
                                         int(self.symmetric), ctypes.c_float(threpq), ctypes.byref(self.c))
        if rc != 0:
            raise RuntimeError("PairList::build_schwarz: basis sets are not symmetric")

    def __del__(self):
        if getattr(self, "_keep", None) is not None:
            return  # arrays owned by numpy (a subsample)
        try:
            lib().oracle_pairlist_free(ctypes.byref(self.c))
        except Exception:
            pass

    def subsample(self, stride):
        """A pair list holding every `stride`-th pair of each class (order and bounds kept), used
        to time the oracle on a bounded sample of a workload."""
        out = OPairList.__new__(OPairList)
        out.b1, out.b2, out.symmetric, out.threpq = self.b1, self.b2, self.symmetric, self.threpq
        p1s, p2s, bds, starts, L1s, L2s = [], [], [], [0], [], []
        for (L1, L2, p1, p2, b) in self.classes():
            p1s.append(p1[::stride])
            p2s.append(p2[::stride])
            bds.append(b[::stride])
            starts.append(starts[-1] + len(p1s[-1]))
            L1s.append(L1)
            L2s.append(L2)
        p1 = np.ascontiguousarray(np.concatenate(p1s) if p1s else np.zeros(0), dtype=np.int32)
        p2 = np.ascontiguousarray(np.concatenate(p2s) if p2s else np.zeros(0), dtype=np.int32)
        bd = np.ascontiguousarray(np.concatenate(bds) if bds else np.zeros(0), dtype=np.float32)
        out._keep = (p1, p2, bd)
        c = _CPairList()
        c.symmetric = int(self.symmetric)
        c.npair = int(p1.size)
        c.p1 = p1.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.p2 = p2.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        c.bound = bd.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        c.ncls = len(L1s)
        for k in range(len(L1s)):
            c.cls_L1[k] = L1s[k]
            c.cls_L2[k] = L2s[k]
            c.cls_start[k] = starts[k]
        c.cls_start[len(L1s)] = starts[-1]
        out.c = c
        return out

    def classes(self):
        """[(L1, L2, p1[], p2[], bound[]), ...] in list order."""
        out = []
        n = self.c.npair
        p1 = np.ctypeslib.as_array(self.c.p1, shape=(max(n, 1),))[:n].copy()
        p2 = np.ctypeslib.as_array(self.c.p2, shape=(max(n, 1),))[:n].copy()
        bd = np.ctypeslib.as_array(self.c.bound, shape=(max(n, 1),))[:n].copy()
        for k in range(self.c.ncls):
            s, e = self.c.cls_start[k], self.c.cls_start[k + 1]
            out.append((self.c.cls_L1[k], self.c.cls_L2[k], p1[s:e], p2[s:e], bd[s:e]))
        return out


def _ewald_arrays(ewald):
    scales, omegas = ewald
    return (np.ascontiguousarray(scales, dtype=np.float64), np.ascontiguousarray(omegas, dtype=np.float64))


def coulomb(ewald, pl12, pl34, D34, thresp, J12=None, counts=None):
    """IntBox::coulombCPU (coulomb.cpp:14-196): J12 += (12|O|34) D34, in each basis's own
    representation.  ewald = (scales, omegas).  Returns J12 (nao1 x nao2)."""
    b1, b2, b3, b4 = pl12.b1, pl12.b2, pl34.b1, pl34.b2
    D34 = np.asarray(D34, dtype=np.float64)
    if D34.shape != (b3.nao, b4.nao):
        raise RuntimeError("Tensor: shape error")
    Dc = np.ascontiguousarray(b3.transform().T @ D34 @ b4.transform())
    Jc = np.zeros((b1.ncart, b2.ncart))
    sc, om = _ewald_arrays(ewald)
    cnt = np.zeros(625, dtype=np.int64)
    lib().oracle_coulomb(ctypes.byref(b1.cstruct), ctypes.byref(b2.cstruct), ctypes.byref(b3.cstruct),
                         ctypes.byref(b4.cstruct), ctypes.byref(pl12.c), ctypes.byref(pl34.c),
                         Dc.ctypes.data, len(sc), sc.ctypes.data, om.ctypes.data, ctypes.c_float(thresp),
                         Jc.ctypes.data, cnt.ctypes.data)
    if counts is not None:
        counts += cnt
    J = b1.transform() @ Jc @ b2.transform().T
    if J12 is not None:
        J12 = np.asarray(J12)
        if J12.shape != (b1.nao, b2.nao):
            raise RuntimeError("Tensor: shape error")
        J = J12 + J
    return J


def exchange(ewald, pl12, pl34, D24, D24symm, thresp, K13=None, counts=None):
    """IntBox::exchangeCPU (exchange.cpp:17-355): K13 += (12|O|34) D24."""
    b1, b2, b3, b4 = pl12.b1, pl12.b2, pl34.b1, pl34.b2
    D24 = np.asarray(D24, dtype=np.float64)
    if D24.shape != (b2.nao, b4.nao):
        raise RuntimeError("Tensor: shape error")
    Dc = np.ascontiguousarray(b2.transform().T @ D24 @ b4.transform())
    if D24symm and not b2.equivalent(b4):
        raise RuntimeError("IntBox::exchangeCPU: D24symm implied, but basis sets 2 and 4 not the same")
    if D24symm:
        D24symm = b1.equivalent(b3)
    Kc = np.zeros((b1.ncart, b3.ncart))
    sc, om = _ewald_arrays(ewald)
    cnt = np.zeros(625, dtype=np.int64)
    lib().oracle_exchange(ctypes.byref(b1.cstruct), ctypes.byref(b2.cstruct), ctypes.byref(b3.cstruct),
                          ctypes.byref(b4.cstruct), ctypes.byref(pl12.c), ctypes.byref(pl34.c),
                          int(pl12 is pl34), Dc.ctypes.data, int(bool(D24symm)), len(sc), sc.ctypes.data,
                          om.ctypes.data, ctypes.c_float(thresp), Kc.ctypes.data, cnt.ctypes.data)
    if counts is not None:
        counts += cnt
    K = b1.transform() @ Kc @ b3.transform().T
    if K13 is not None:
        K13 = np.asarray(K13)
        if K13.shape != (b1.nao, b3.nao):
            raise RuntimeError("Tensor: shape error")
        K = K13 + K
    return K


def _quartet_call(fn, width, ewald, bases, P):
    P = [np.ascontiguousarray(x, dtype=np.int32).reshape(-1) for x in P]
    n = P[0].size
    out = np.zeros(max(n, 1) * width)
    sc, om = _ewald_arrays(ewald)
    fn(*[ctypes.byref(b.cstruct) for b in bases], n, *[x.ctypes.data for x in P], len(sc), sc.ctypes.data,
       om.ctypes.data, out.ctypes.data)
    return out[:n * width]


def quartet_max(ewald, b1, b2, b3, b4, P1, P2, P3, P4):
    """max over Cartesian components of |(P1 P2 | O | P3 P4)| for each primitive quartet (the
    per-quartet evaluation of exchange.cpp:191-324 without the D contraction)."""
    return _quartet_call(lib().oracle_quartet_max, 1, ewald, (b1, b2, b3, b4), (P1, P2, P3, P4))


def quartet_tensor(ewald, b1, b2, b3, b4, P1, P2, P3, P4):
    """All Cartesian components of each primitive quartet, shape (n, N1, N2, N3, N4)."""
    P = [np.ascontiguousarray(x, dtype=np.int32).reshape(-1) for x in (P1, P2, P3, P4)]
    flat = _quartet_call(lib().oracle_quartet_tensor, 1296, ewald, (b1, b2, b3, b4), P).reshape(-1, 1296)
    out = []
    for k in range(flat.shape[0]):
        N = [(int(b.L[p[k]]) + 1) * (int(b.L[p[k]]) + 2) // 2 for b, p in zip((b1, b2, b3, b4), P)]
        out.append(flat[k, :N[0] * N[1] * N[2] * N[3]].reshape(N))
    return out


def rys(n, T):
    """Rys nodes x = t^2 and weights at T (the table the kernels use)."""
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
    x = np.zeros((T.size, n))
    w = np.zeros((T.size, n))
    lib().oracle_rys(n, T.size, T.ctypes.data, x.ctypes.data, w.ctypes.data)
    return x, w


def _accumulate(b1, b2, Mc, M):
    out = b1.transform() @ Mc @ b2.transform().T
    if M is not None:
        M = np.asarray(M)
        if M.shape != (b1.nao, b2.nao):
            raise RuntimeError("Tensor: shape error")
        out = M + out
    return out


def overlap(pl, S=None):
    """IntBox::overlapCPU (src/cpp/intbox/cpu/overlap.cpp:14-147): S += <a|b>."""
    b1, b2 = pl.b1, pl.b2
    Sc = np.zeros((b1.ncart, b2.ncart))
    lib().oracle_overlap_kinetic(ctypes.byref(b1.cstruct), ctypes.byref(b2.cstruct), ctypes.byref(pl.c),
                                 Sc.ctypes.data, None)
    return _accumulate(b1, b2, Sc, S)


def kinetic(pl, T=None):
    """IntBox::kineticCPU (src/cpp/intbox/cpu/kinetic.cpp:14-175): T += <a|-1/2 nabla^2|b>."""
    b1, b2 = pl.b1, pl.b2
    Tc = np.zeros((b1.ncart, b2.ncart))
    lib().oracle_overlap_kinetic(ctypes.byref(b1.cstruct), ctypes.byref(b2.cstruct), ctypes.byref(pl.c),
                                 None, Tc.ctypes.data)
    return _accumulate(b1, b2, Tc, T)


def potential(ewald, pl, xyzc, V=None):
    """IntBox::potentialCPU (src/cpp/intbox/cpu/potential.cpp:14-147): V += sum_A <a|O(r_A)|b> (-c_A)."""
    b1, b2 = pl.b1, pl.b2
    xyzc = np.ascontiguousarray(xyzc, dtype=np.float64).reshape(-1, 4)
    sc, om = _ewald_arrays(ewald)
    Vc = np.zeros((b1.ncart, b2.ncart))
    lib().oracle_potential(ctypes.byref(b1.cstruct), ctypes.byref(b2.cstruct), ctypes.byref(pl.c), len(sc),
                           sc.ctypes.data, om.ctypes.data, xyzc.shape[0], xyzc.ctypes.data, Vc.ctypes.data)
    return _accumulate(b1, b2, Vc, V)


def eriJ(ewald, pl12, pl34, C1, C2, C3, C4, thresp, same12=False, same34=False, same1234=False):
    """IntBox::eriJ (src/cpp/intbox/mo_eri.cpp:10-157): I[i,j,k,l] = (C1_i C2_j | C3_k C4_l) by one J
    build over (pl34 | pl12) per (i, j) density C1_i C2_j^T, transformed with C3, C4; the flags stand
    for the reference's pointer tests C1 == C2, C3 == C4, (C1, C2) == (C3, C4)."""
    C1, C2, C3, C4 = (np.asarray(c, dtype=np.float64) for c in (C1, C2, C3, C4))
    n1, n2, n3, n4 = C1.shape[1], C2.shape[1], C3.shape[1], C4.shape[1]
    I = np.zeros((n1, n2, n3, n4))
    for i in range(n1):
        for j in range(n2):
            if same12 and i < j:
                continue
            D = np.outer(C1[:, i], C2[:, j])
            J = coulomb(ewald, pl34, pl12, D, thresp)
            I[i, j] = C3.T @ J @ C4
            if same12 and i != j:
                I[j, i] = I[i, j]
    if same12:
        for p in range(n1):
            for q in range(p):
                I[p, q] = I[q, p] = 0.5 * (I[p, q] + I[q, p])
    if same34:
        I = 0.5 * (I + I.transpose(0, 1, 3, 2))
    if same1234:
        M = I.reshape(n1 * n2, n3 * n4)
        I = (0.5 * (M + M.T)).reshape(n1, n2, n3, n4)
    return I
