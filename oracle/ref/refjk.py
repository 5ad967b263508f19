"""TEST INFRASTRUCTURE ONLY -- runs the reference's own CPU J/K path (built from its sources by
oracle/ref/Makefile into oracle/_ref/ref_jk) on a problem written from any basis object exposing
`.primitives` (fields c, e, x, y, z, c0, L, is_pure, aoIdx, cartIdx, primIdx, shellIdx, atomIdx).

The call is the reference's IntBox::coulombCPU / IntBox::exchangeCPU
(src/cpp/intbox/cpu/coulomb.cpp:14-196, src/cpp/intbox/cpu/exchange.cpp:17-355) on
PairList::build_schwarz(basis, basis, symmetric, threpq) (src/cpp/core/pair_list.cpp), with the
bare Coulomb operator (Ewald scales [1], omegas [-1]).
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BINARY = os.path.normpath(os.path.join(_HERE, "..", "_ref", "ref_jk"))


def available():
    return os.path.exists(BINARY)


def build():
    subprocess.check_call(["make", "-s", "-j8", "-C", _HERE])


def run(basis, D, threpq=1e-14, thresp=1e-14, symmetric=True, D24symm=True, nrep=1, J=True, K=True,
        nthreads=None, stride=1):
    """Returns (J, K, timings) with timings = (median J s, median K s, pair-list s).  With
    stride > 1 the bra list keeps every stride-th pair of each class (the oracle's
    OPairList.subsample) and the ket list is whole."""
    if not available():
        raise RuntimeError("reference build missing: run `make -C oracle/ref`")
    prims = list(basis.primitives)
    D = np.ascontiguousarray(D, dtype=np.float64)
    nao = D.shape[0]
    buf = [struct.pack("<i", len(prims))]
    for p in prims:
        buf.append(struct.pack("<6d7i", p.c, p.e, p.x, p.y, p.z, p.c0, p.L, int(p.is_pure), p.aoIdx,
                               p.cartIdx, p.primIdx, p.shellIdx, p.atomIdx))
    buf.append(struct.pack("<i", nao))
    buf.append(D.tobytes())
    buf.append(struct.pack("<2f6i", threpq, thresp, int(symmetric), int(D24symm), nrep, int(J), int(K), stride))
    env = dict(os.environ)
    if nthreads is not None:
        env["OMP_NUM_THREADS"] = str(nthreads)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(b"".join(buf))
        subprocess.check_call([BINARY, fin, fout], env=env)
        out = np.fromfile(fout, dtype=np.float64)
    n2 = nao * nao
    return out[:n2].reshape(nao, nao), out[n2:2 * n2].reshape(nao, nao), tuple(out[2 * n2:2 * n2 + 3])
