"""Writes tests/golden/ref_jk_*.npz: J and K of the reference's own CPU path, built from its
sources by oracle/ref/Makefile (oracle/_ref/ref_jk), on small seeded workloads.

Each file holds the inputs that fix the problem (molecule name, basis, seed, thresholds, D) and
the reference outputs J, K.  Run in a container that has the reference sources:
    make -C oracle/ref && python tests/golden/make_ref_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))

import lightspeed_amd as ls  # noqa: E402
from lightspeed_amd import workloads  # noqa: E402
import refjk  # noqa: E402

# name: (cluster size, box (A), cluster seed, basis, nocc, density seed, threpq, thresp)
CASES = {
    "h2o10_631g": (10, 10.0, 10, "6-31g", 50, 0, 1e-14, 1e-12),
    "h2o4_ccpvdz": (4, 7.0, 4, "cc-pvdz", 20, 1, 1e-14, 1e-12),
    "h2o3_631gss": (3, 6.0, 3, "6-31gss", 15, 2, 1e-14, 1e-13),
}


def problem(name):
    n, box, seed, bname, nocc, dseed, threpq, thresp = CASES[name]
    mol = workloads.water_cluster(n, box, seed)
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = workloads.random_density(bas.nao, nocc, dseed)
    return bas, D, threpq, thresp


def main():
    for name in CASES:
        bas, D, threpq, thresp = problem(name)
        J, K, _ = refjk.run(bas, D, threpq, thresp, symmetric=True, D24symm=True)
        np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_jk_%s.npz" % name),
                            D=D, J=J, K=K, threpq=threpq, thresp=thresp, nwater=CASES[name][0],
                            box=CASES[name][1], seed=CASES[name][2], basis=CASES[name][3])
        print(name, bas.nao, np.abs(J).max(), np.abs(K).max())


if __name__ == "__main__":
    main()
