"""ctypes binding of the C-ABI (include/lsjk.h), the way a caller without the C++ layer binds it.

Every function used here gets its argtypes and restype, so 64-bit pointers are passed whole.
Array dtypes: float64 for c, e, xyz (nprim x 3), densities, scales and omegas; int32 for L, cart,
p1, p2; float32 for the pair bounds.  Handles and device pointers are c_void_p.
"""
import ctypes

import numpy as np

import lightspeed_amd as ls

vp = ctypes.c_void_p
ci = ctypes.c_int
cf = ctypes.c_float
csz = ctypes.c_size_t


def load():
    L = ctypes.CDLL(ls.NATIVE_LIBRARY)
    L.lsjk_last_error.restype = ctypes.c_char_p
    L.lsjk_last_error.argtypes = []
    L.lsjk_init.argtypes = [ci]
    L.lsjk_basis_create.argtypes = [ci, vp, vp, vp, vp, vp, ci, ctypes.POINTER(vp)]
    L.lsjk_basis_destroy.argtypes = [vp]
    L.lsjk_pairlist_create.argtypes = [vp, vp, ci, ci, vp, vp, vp, ctypes.POINTER(vp)]
    L.lsjk_pairlist_destroy.argtypes = [vp]
    build = [vp, vp, vp, ci, vp, vp, cf, cf, vp]
    L.lsjk_coulomb.argtypes = build
    L.lsjk_coulomb_noalloc.argtypes = build
    L.lsjk_exchange.argtypes = [vp, vp, vp, ci, ci, vp, vp, cf, cf, vp]
    L.lsjk_malloc.argtypes = [ctypes.POINTER(vp), csz]
    L.lsjk_free.argtypes = [vp]
    L.lsjk_memcpy_htod.argtypes = [vp, vp, csz]
    L.lsjk_memcpy_dtoh.argtypes = [vp, vp, csz]
    L.lsjk_memset_zero.argtypes = [vp, csz]
    for f in (L.lsjk_init, L.lsjk_basis_create, L.lsjk_basis_destroy, L.lsjk_pairlist_create,
              L.lsjk_pairlist_destroy, L.lsjk_coulomb, L.lsjk_coulomb_noalloc, L.lsjk_exchange, L.lsjk_malloc,
              L.lsjk_free, L.lsjk_memcpy_htod, L.lsjk_memcpy_dtoh, L.lsjk_memset_zero):
        f.restype = ci
    return L


class Lib:
    def __init__(self, device=0):
        self.L = load()
        self.check(self.L.lsjk_init(device))

    def check(self, rc):
        if rc:
            raise RuntimeError(self.L.lsjk_last_error().decode())

    def basis(self, bas):
        """lsjk_basis of a lightspeed_amd.Basis (primitive arrays in its own order)."""
        prims = bas.primitives
        c = np.array([p.c for p in prims], np.float64)
        e = np.array([p.e for p in prims], np.float64)
        xyz = np.array([[p.x, p.y, p.z] for p in prims], np.float64)
        Ls = np.array([p.L for p in prims], np.int32)
        cart = np.array([p.cartIdx for p in prims], np.int32)
        h = vp()
        self.check(self.L.lsjk_basis_create(len(prims), c.ctypes.data, e.ctypes.data, xyz.ctypes.data,
                                            Ls.ctypes.data, cart.ctypes.data, bas.ncart, ctypes.byref(h)))
        return h

    def pairlist(self, hb1, hb2, pl):
        """lsjk_pairlist of a lightspeed_amd.PairList (its classes in order, bounds descending)."""
        p1, p2, b = [], [], []
        for l in pl.pairlists:
            a1, a2, ab = l.arrays()
            p1.append(a1)
            p2.append(a2)
            b.append(ab)
        p1 = np.ascontiguousarray(np.concatenate(p1), np.int32)
        p2 = np.ascontiguousarray(np.concatenate(p2), np.int32)
        b = np.ascontiguousarray(np.concatenate(b), np.float32)
        h = vp()
        self.check(self.L.lsjk_pairlist_create(hb1, hb2, 1 if pl.is_symmetric else 0, len(p1), p1.ctypes.data,
                                               p2.ctypes.data, b.ctypes.data, ctypes.byref(h)))
        return h

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        d = vp()
        self.check(self.L.lsjk_malloc(ctypes.byref(d), a.nbytes))
        self.check(self.L.lsjk_memcpy_htod(d, a.ctypes.data, a.nbytes))
        return d

    def zeros(self, shape):
        n = int(np.prod(shape)) * 8
        d = vp()
        self.check(self.L.lsjk_malloc(ctypes.byref(d), n))
        self.check(self.L.lsjk_memset_zero(d, n))
        return d

    def download(self, d, shape):
        out = np.zeros(shape, np.float64)
        self.check(self.L.lsjk_memcpy_dtoh(out.ctypes.data, d, out.nbytes))
        return out

    def free(self, d):
        self.check(self.L.lsjk_free(d))

    @staticmethod
    def ewald(scales, omegas):
        return np.array(scales, np.float64), np.array(omegas, np.float64)

    def coulomb(self, pl12, pl34, dD, ew, thresp, thredp, dJ):
        s, o = ew
        self.check(self.L.lsjk_coulomb(pl12, pl34, dD, len(s), s.ctypes.data, o.ctypes.data, thresp, thredp, dJ))

    def coulomb_noalloc(self, pl12, pl34, dD, ew, thresp, thredp, dJ):
        s, o = ew
        self.check(self.L.lsjk_coulomb_noalloc(pl12, pl34, dD, len(s), s.ctypes.data, o.ctypes.data, thresp, thredp,
                                               dJ))

    def exchange(self, pl12, pl34, dD, D24symm, ew, thresp, thredp, dK):
        s, o = ew
        self.check(self.L.lsjk_exchange(pl12, pl34, dD, 1 if D24symm else 0, len(s), s.ctypes.data, o.ctypes.data,
                                        thresp, thredp, dK))
