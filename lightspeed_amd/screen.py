"""The learned quartet screen: a small network, trained on the GPU on exchange quartet magnitudes
that the exact Rys kernel generates, that tightens the Schwarz estimate per exchange work unit.

Network: ni inputs -> 3 hidden layers of 64 ReLU units -> one linear output, FP32, evaluated and
trained by the library's own HIP kernels (lsjk_mlp_forward / lsjk_mlp_grad / lsjk_mlp_adam).  This
module holds the device buffers and the training loop; the arithmetic is on the GPU.
"""
import ctypes

import numpy as np

from . import NATIVE_LIBRARY, hip_init

_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cf = ctypes.c_float
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(NATIVE_LIBRARY)
        L.lsjk_last_error.restype = ctypes.c_char_p
        L.lsjk_malloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        L.lsjk_free.argtypes = [_vp]
        L.lsjk_memcpy_htod.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.lsjk_memcpy_dtoh.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.lsjk_memset_zero.argtypes = [_vp, ctypes.c_size_t]
        L.lsjk_mlp_nparam.argtypes = [_ci, ctypes.POINTER(_ci)]
        L.lsjk_mlp_forward.argtypes = [_ci, _vp, _ci, _vp, _vp]
        L.lsjk_mlp_grad.argtypes = [_ci, _vp, _ci, _vp, _vp, _vp, _vp, _cf, _vp, _vp]
        L.lsjk_mlp_adam.argtypes = [_ci, _vp, _vp, _vp, _vp, _cf, _cf, _cf, _cf, _ci]
        _lib = L
    return _lib


def check(rc):
    if rc:
        raise RuntimeError("lightspeed_amd HIP backend: " + lib().lsjk_last_error().decode())


class DevBuf:
    """A raw device buffer of a numpy dtype (the library's allocator)."""

    def __init__(self, n, dtype):
        hip_init(0)
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(1, self.n) * self.dtype.itemsize
        self.ptr = _vp()
        check(lib().lsjk_malloc(ctypes.byref(self.ptr), self.nbytes))
        check(lib().lsjk_memset_zero(self.ptr, self.nbytes))

    @classmethod
    def of(cls, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        b = cls(a.size, dtype)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError("DevBuf.upload: size mismatch")
        if a.size:
            check(lib().lsjk_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def download(self):
        out = np.zeros(self.n, dtype=self.dtype)
        if self.n:
            check(lib().lsjk_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        check(lib().lsjk_memset_zero(self.ptr, self.nbytes))

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib().lsjk_free(self.ptr)
            self.ptr = _vp()


def nparam(ni):
    n = _ci()
    check(lib().lsjk_mlp_nparam(ni, ctypes.byref(n)))
    return n.value


class MLP:
    """The screen network on the device (FP32 parameters)."""

    HIDDEN = 64

    def __init__(self, ni, params):
        self.ni = int(ni)
        self.np = nparam(self.ni)
        params = np.asarray(params, dtype=np.float32).reshape(-1)
        if params.size != self.np:
            raise ValueError("MLP: %d parameters expected" % self.np)
        self.p = DevBuf.of(params, np.float32)

    @classmethod
    def init(cls, ni, seed):
        """He-normal weights, zero biases."""
        rng = np.random.default_rng(seed)
        H = cls.HIDDEN
        p = np.zeros(nparam(ni), dtype=np.float64)
        k = 0
        for size, fan in ((H * ni, ni), (H, 0), (H * H, H), (H, 0), (H * H, H), (H, 0), (H, H), (1, 0)):
            if fan:
                p[k:k + size] = rng.standard_normal(size) * np.sqrt(2.0 / fan)
            k += size
        return cls(ni, p)

    def params(self):
        return self.p.download()

    def forward_dev(self, X, n, out):
        check(lib().lsjk_mlp_forward(self.ni, self.p.ptr, int(n), X.ptr, out.ptr))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.ni)
        dX = DevBuf.of(X, np.float32)
        out = DevBuf(len(X), np.float32)
        self.forward_dev(dX, len(X), out)
        return out.download()

    def grad(self, X, y, idx=None, sw=None, wu=1.0):
        """(loss, gradient) of the asymmetric squared error over rows idx of (X, y) (all rows when
        idx is None), computed by the HIP backward kernel."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.ni)
        dX = DevBuf.of(X, np.float32)
        dy = DevBuf.of(y, np.float32)
        dsw = DevBuf.of(sw, np.float32) if sw is not None else None
        didx = DevBuf.of(idx, np.int32) if idx is not None else None
        nb = len(idx) if idx is not None else len(X)
        g = DevBuf(self.np, np.float32)
        loss = DevBuf(1, np.float32)
        check(lib().lsjk_mlp_grad(self.ni, self.p.ptr, nb, didx.ptr if didx else None, dX.ptr, dy.ptr,
                                  dsw.ptr if dsw else None, float(wu), g.ptr, loss.ptr))
        return float(loss.download()[0]), g.download()

    def train(self, X, y, epochs=40, batch=4096, lr=2.0e-3, wu=4.0, sw=None, seed=0, log=None):
        """Adam on mini-batches of shuffled rows; everything but the shuffling runs on the GPU.
        Returns the mean training loss of each epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.ni)
        n = len(X)
        dX = DevBuf.of(X, np.float32)
        dy = DevBuf.of(y, np.float32)
        dsw = DevBuf.of(sw, np.float32) if sw is not None else None
        g = DevBuf(self.np, np.float32)
        m = DevBuf(self.np, np.float32)
        v = DevBuf(self.np, np.float32)
        loss = DevBuf(1, np.float32)
        didx = DevBuf(n, np.int32)
        rng = np.random.default_rng(seed)
        t = 0
        history = []
        for ep in range(epochs):
            perm = rng.permutation(n).astype(np.int32)
            didx.upload(perm)
            loss.zero()
            nstep = 0
            for b0 in range(0, n, batch):
                nb = min(batch, n - b0)
                g.zero()
                sub = ctypes.c_void_p(didx.ptr.value + 4 * b0)
                check(lib().lsjk_mlp_grad(self.ni, self.p.ptr, nb, sub, dX.ptr, dy.ptr,
                                          dsw.ptr if dsw else None, float(wu), g.ptr, loss.ptr))
                t += 1
                check(lib().lsjk_mlp_adam(self.np, self.p.ptr, g.ptr, m.ptr, v.ptr, float(lr), 0.9, 0.999, 1.0e-8, t))
                nstep += 1
            history.append(float(loss.download()[0]) / max(nstep, 1))
            if log:
                log("epoch %d loss %.4g" % (ep, history[-1]))
        return history


# ------------------------------------------------------------------ unit data, training, certification

NFEAT = 16
SP_CLASSES = [(a, b, c, d) for a in (0, 1) for b in (0, 1) for c in (0, 1) for d in (0, 1)]
FEATURES = ("L1", "L2", "L3", "L4", "ln e(bra row)", "ln min e(bra partners)", "ln min e(ket rows)",
            "ln min e(ket partners)", "R(bra row, ket centroid)", "ket group radius", "max bra extent",
            "max ket extent", "ln max bra bound", "ln max ket bound", "ln ss ratio (closest approach)",
            "ln ss ratio (centroid distance)")


def _bind_screen():
    L = lib()
    if getattr(L, "_screen_bound", False):
        return L
    L.lsjk_unit_features.argtypes = [_vp, _ci, _ci, _ci, _ci, _cf, ctypes.POINTER(_ci), _vp]
    L.lsjk_unit_magnitudes.argtypes = [_vp, _ci, _ci, _ci, _ci, _cf, _cf, _vp]
    L.lsjk_screen_create.argtypes = [_ci, _vp, _vp, _vp, _ci, _vp, _vp, _cf, ctypes.POINTER(_vp)]
    L.lsjk_screen_destroy.argtypes = [_vp]
    L.lsjk_exchange_screened.argtypes = [_vp, _vp, _vp, _ci, _ci, _vp, _vp, _cf, _cf, _vp, _vp]
    L._screen_bound = True
    return L


def class_index(L1, L2, L3, L4):
    return ((L1 * 5 + L2) * 5 + L3) * 5 + L4


def unit_data(pl, thresp=1.0e-12, smin=1.0e-14, classes=SP_CLASSES):
    """Features (n, NFEAT), targets ln(max_q M_q / S_q) (n,) and class index (n,) of the exchange
    work units of the symmetric list pl on both sides, for the given classes.  The targets come from
    the exact Rys evaluation on the GPU; units without a quartet of Schwarz product >= smin are
    dropped."""
    L = _bind_screen()
    ptr = _vp(pl.device_ptr)
    Xs, ys, cs = [], [], []
    for cls in classes:
        n = _ci()
        check(L.lsjk_unit_features(ptr, *cls, float(thresp), ctypes.byref(n), None))
        if n.value == 0:
            continue
        feat = np.zeros((n.value, NFEAT), dtype=np.float32)
        check(L.lsjk_unit_features(ptr, *cls, float(thresp), ctypes.byref(n), feat.ctypes.data))
        mag = np.zeros(n.value, dtype=np.float32)
        check(L.lsjk_unit_magnitudes(ptr, *cls, float(thresp), float(smin), mag.ctypes.data))
        keep = mag > 0
        Xs.append(feat[keep])
        ys.append(np.log(mag[keep].astype(np.float64)))
        cs.append(np.full(int(keep.sum()), class_index(*cls), dtype=np.int32))
    if not Xs:
        return np.zeros((0, NFEAT), np.float32), np.zeros(0), np.zeros(0, np.int32)
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(cs)


def training_clusters(sizes=(8, 12, 16, 20, 24, 28, 32), seed0=100):
    """Seeded water clusters of the training set: (H2O)_n, O atoms in a cube of edge 20 A (n/64)^(1/3)
    (the density of the (H2O)64 / 20 A workload), cluster seed seed0 + n."""
    from . import workloads as W
    return [W.water_cluster(n, 20.0 * (n / 64.0) ** (1.0 / 3.0), seed0 + n) for n in sizes]


def training_set(basis="6-31gss", sizes=(8, 12, 16, 20, 24, 28, 32), seed0=100, thresp=1.0e-12, smin=1.0e-14,
                 threpq=1.0e-12, log=None):
    """Unit data of every training cluster (see training_clusters)."""
    from . import Basis, PairList
    Xs, ys, cs = [], [], []
    for mol in training_clusters(sizes, seed0):
        bas = Basis.from_gbs_file(mol, basis)
        pl = PairList.build_schwarz(bas, bas, True, threpq)
        X, y, c = unit_data(pl, thresp, smin)
        if log:
            log("%s: %d units" % (mol.name, len(y)))
        Xs.append(X)
        ys.append(y)
        cs.append(c)
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(cs)


class Screen:
    """A trained screen: network parameters, input normalisation, per class and bin of the network
    output a certified margin (delta (625, nbin), NaN: not certified; bins split at fedge
    (625, nbin - 1)) and the smallest Schwarz product the margins hold for."""

    def __init__(self, params, mean, scale, delta, smin, report=None, fedge=None):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.mean = np.ascontiguousarray(mean, dtype=np.float32)
        self.scale = np.ascontiguousarray(scale, dtype=np.float32)
        self.delta = np.ascontiguousarray(np.asarray(delta, dtype=np.float64).reshape(625, -1))
        self.nbin = self.delta.shape[1]
        self.fedge = np.ascontiguousarray(np.zeros((625, max(self.nbin - 1, 1)), np.float32) if fedge is None else
                                          np.asarray(fedge, dtype=np.float32).reshape(625, max(self.nbin - 1, 1)))
        self.ni = len(self.mean)
        self.smin = float(smin)
        self.report = report or {}
        self._h = None

    @property
    def handle(self):
        if self._h is None:
            L = _bind_screen()
            hip_init(0)
            h = _vp()
            check(L.lsjk_screen_create(self.ni, self.params.ctypes.data, self.mean.ctypes.data, self.scale.ctypes.data,
                                       self.nbin, self.fedge.ctypes.data, self.delta.ctypes.data, self.smin,
                                       ctypes.byref(h)))
            self._h = h
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().lsjk_screen_destroy(self._h)
            self._h = None

    def certified(self):
        return [c for c in range(625) if np.isfinite(self.delta[c]).any()]

    def predict(self, X, cls):
        """tau = min(1, exp(f(x) + delta[class, bin of f])) per unit (1 where not certified)."""
        net = MLP(self.ni, self.params)
        X = np.asarray(X, dtype=np.float32)[:, :self.ni]
        f = net.forward((X - self.mean) / self.scale).astype(np.float64)
        cls = np.asarray(cls)
        b = np.zeros(len(f), dtype=np.int64)
        for k in range(self.nbin - 1):
            b += f >= self.fedge[cls, k]
        d = self.delta[cls, b]
        return np.where(np.isfinite(d), np.minimum(1.0, np.exp(f + np.nan_to_num(d, nan=0.0))), 1.0)

    def exchange_device(self, ewald, pl12, pl34, D_ptr, D24symm, thresp, thredp, K_ptr):
        """lsjk_exchange_screened on HBM-resident Cartesian D24 / K13 (accumulated)."""
        L = _bind_screen()
        s = np.ascontiguousarray(ewald.scales, dtype=np.float64)
        o = np.ascontiguousarray(ewald.omegas, dtype=np.float64)
        check(L.lsjk_exchange_screened(_vp(pl12.device_ptr), _vp(pl34.device_ptr), _vp(D_ptr), int(bool(D24symm)),
                                       len(s), s.ctypes.data, o.ctypes.data, float(thresp), float(thredp), self.handle,
                                       _vp(K_ptr)))

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, scale=self.scale, delta=self.delta, fedge=self.fedge,
                 smin=np.array([self.smin]), report_keys=np.array(sorted(self.report)),
                 report_vals=np.array([float(self.report[k]) for k in sorted(self.report)]))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        rep = dict(zip([str(k) for k in z["report_keys"]], [float(v) for v in z["report_vals"]]))
        return cls(z["params"], z["mean"], z["scale"], z["delta"], float(z["smin"][0]), rep,
                   z["fedge"] if "fedge" in z.files else None)


def train_screen(X, y, cls, smin=1.0e-14, epochs=30, batch=4096, lr=2.0e-3, wu=4.0, slack=0.05, nbin=8, min_bin=25,
                 min_coverage=0.999, seed=0, log=None):
    """Train the network on 60 % of the units, calibrate margins on 20 % and test them on the last
    20 %.  Per class, the calibration units are split into nbin equal-count bins of the network
    output; a bin's margin is its largest residual y - f plus `slack`, so that units predicted far
    (small f, where the factor matters) get their own margin instead of the class's worst case.
    A bin is certified when it holds at least `min_bin` calibration units, and a class only when the
    test split's coverage (units with y <= f + delta of f's bin) is at least `min_coverage`.  Returns
    the Screen; its report holds the training and per-class numbers."""
    X = np.asarray(X, dtype=np.float32)
    y = np.asarray(y, dtype=np.float64)
    cls = np.asarray(cls)
    ni = X.shape[1]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    n1, n2 = int(0.6 * len(y)), int(0.8 * len(y))
    tr, ca, te = perm[:n1], perm[n1:n2], perm[n2:]
    mean = X[tr].mean(0)
    scale = np.maximum(X[tr].std(0), 1.0e-3).astype(np.float32)
    Xn = ((X - mean) / scale).astype(np.float32)
    net = MLP.init(ni, seed + 1)
    hist = net.train(Xn[tr], y[tr].astype(np.float32), epochs=epochs, batch=batch, lr=lr, wu=wu, seed=seed + 2,
                     log=log)
    f = net.forward(Xn).astype(np.float64)
    delta = np.full((625, nbin), np.nan)
    fedge = np.zeros((625, max(nbin - 1, 1)), np.float32)
    report = {"n_train": len(tr), "n_cal": len(ca), "n_test": len(te), "loss_first": hist[0], "loss_last": hist[-1],
              "nbin": nbin}
    for c in np.unique(cls):
        mc, mt = ca[cls[ca] == c], te[cls[te] == c]
        if len(mc) < min_bin * nbin or len(mt) == 0:
            continue
        edges = np.quantile(f[mc], np.linspace(0, 1, nbin + 1)[1:-1]).astype(np.float32)
        fedge[c, :nbin - 1] = edges
        bc = np.searchsorted(edges, f[mc].astype(np.float32), side="right")
        d = np.full(nbin, np.nan)
        for k in range(nbin):
            r = y[mc][bc == k] - f[mc][bc == k]
            if len(r) >= min_bin:
                d[k] = float(np.max(r)) + slack
        bt = np.searchsorted(edges, f[mt].astype(np.float32), side="right")
        dt = d[bt]
        ok = np.isfinite(dt)
        cov = float(np.mean(np.where(ok, y[mt] <= f[mt] + np.nan_to_num(dt), True)))
        tight = float(np.mean(np.where(ok, np.minimum(1.0, np.exp(f[mt] + np.nan_to_num(dt))), 1.0)))
        report["coverage_%d" % c] = cov
        report["mean_tau_%d" % c] = tight
        if cov >= min_coverage:
            delta[c] = d
        if log:
            log("class %d: cal %d test %d delta %s coverage %.5f mean tau %.3f%s" % (
                c, len(mc), len(mt), np.array2string(d, precision=2), cov, tight,
                "" if cov >= min_coverage else " (not certified)"))
    return Screen(net.params(), mean, scale, delta, smin, report, fedge)


DEFAULT_MODEL = __import__("os").path.join(__import__("os").path.dirname(__file__), "data", "screen_h2o_631gss.npz")


def main(argv=None):
    """Generate the training set on the GPU, train and certify the screen, save it.

    python -m lightspeed_amd.screen [--out PATH] [--sizes 8,12,...] [--epochs N]
    """
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--basis", default="6-31gss")
    ap.add_argument("--sizes", default="8,12,16,20,24,28,32")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)

    def log(m):
        print("[screen] " + m, file=sys.stderr, flush=True)

    t0 = time.time()
    X, y, c = training_set(a.basis, tuple(int(v) for v in a.sizes.split(",")), log=log)
    t1 = time.time()
    log("training set: %d units in %.1f s" % (len(y), t1 - t0))
    scr = train_screen(X, y, c, epochs=a.epochs, seed=a.seed, log=log)
    t2 = time.time()
    scr.report["data_s"] = t1 - t0
    scr.report["train_s"] = t2 - t1
    scr.save(a.out)
    log("trained in %.1f s; certified classes %s; saved %s" % (t2 - t1, scr.certified(), a.out))
    print(json.dumps({k: float(v) for k, v in scr.report.items()}))


if __name__ == "__main__":
    main()
