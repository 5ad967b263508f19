#!/usr/bin/env python3
"""J+K build benchmark on one MI355X.

One step = one direct Coulomb build plus one direct exchange build (IntBox::coulomb +
IntBox::exchange, symmetric Schwarz pair list, D24symm) over an HBM-resident density, on the
workload the target is quoted on: (H2O)64 random cluster (20 A cube, seed 64), 6-31G** (1600
functions), Schwarz cutoff 1e-12, thresp = 1e-12, D = C C^T from a seeded random orthonormal C
(nocc = 320).  Route: plain FP64.

value = ERI quartets/s = (N_J + N_K) / t_JK with N counted under the reference's screening
conventions (the plain FP64 route's quartets), t_JK the wall time of one J+K step.
roofline: the dominant quartet kernel (largest mean launch time over the timed steps, HIP events
on the library stream): algorithmic FLOPs per launch = sum over its classes of
N(class) x F(class), F = 40 + nrys (30 + 9 G + 3 Pi) + 2 Pi (Pi = N1 N2 N3 N4,
G = (L1+L2+1)(L3+L4+1), nrys = (L1+L2+L3+L4)/2 + 1), against the FP64 peak 78.6 TFLOP/s.
cpu_baseline: the reference's own CPU IntBox::coulombCPU + IntBox::exchangeCPU, compiled from
its sources into oracle/_ref/ref_jk, timed on a bounded sample of the same workload (the bra list
keeps every `--cpu-stride`-th pair of each class, the ket list is whole), all threads of the host.
The sample's quartet count comes from the CPU FP64 oracle run on the same sample (its counts
equal the reference's evaluated quartets); where the reference build is absent the oracle's own
time is reported (kind "port").

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload h2o64_631gss]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "J+K build time (s) & ERI quartets/s, mixed vs FP64 GPU at |ΔE|≤1e-6 Eh"
FP64_PEAK = 78.6e12


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ncart(L):
    return (L + 1) * (L + 2) // 2


def flops_per_quartet(L1, L2, L3, L4):
    L = L1 + L2 + L3 + L4
    nrys = L // 2 + 1
    Pi = ncart(L1) * ncart(L2) * ncart(L3) * ncart(L4)
    G = (L1 + L2 + 1) * (L3 + L4 + 1)
    return 40 + nrys * (30 + 9 * G + 3 * Pi) + 2 * Pi


def cls_of(idx):
    L4 = idx % 5
    L3 = (idx // 5) % 5
    L2 = (idx // 25) % 5
    L1 = idx // 125
    return L1, L2, L3, L4


def cpu_baseline(bas, D, stride, thresp):
    """Time the reference build (or the oracle) on a sample; the oracle supplies the count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    import oracle as O
    import refjk
    ob = O.OBasis(bas)
    opl = O.OPairList(ob, ob, True, 1.0e-12)
    sub = opl.subsample(stride)
    ew = ([1.0], [-1.0])
    cj = np.zeros(625, np.int64)
    ck = np.zeros(625, np.int64)
    t0 = time.perf_counter()
    O.coulomb(ew, sub, opl, D, thresp, counts=cj)
    t1 = time.perf_counter()
    O.exchange(ew, sub, opl, D, True, thresp, counts=ck)
    t2 = time.perf_counter()
    n = int(cj.sum() + ck.sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    port = "CPU FP64 oracle (OpenMP, %d threads): J %.2f s + K %.2f s" % (threads, t1 - t0, t2 - t1)
    sample = "every %d-th bra pair of each class of the same workload, whole ket list: %d J + %d K quartets" % (
        stride, int(cj.sum()), int(ck.sum()))
    if refjk.available():
        _, _, (tj, tk, _) = refjk.run(bas, D, 1.0e-12, thresp, symmetric=True, D24symm=True, stride=stride,
                                      nthreads=threads)
        return {"value": n / (tj + tk), "unit": "ERI quartets/s", "cores": threads, "kind": "reference",
                "sample": "reference IntBox::coulombCPU + exchangeCPU built from its sources (OpenMP, %d "
                          "threads) on %s in J %.2f s + K %.2f s; quartets counted by the oracle on the same "
                          "sample (%s)" % (threads, sample, tj, tk, port)}
    return {"value": n / (t2 - t0), "unit": "ERI quartets/s", "cores": threads, "kind": "port",
            "sample": "%s on %s" % (port, sample)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o64_631gss")
    ap.add_argument("--cpu-stride", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="short run for hardware-counter passes: one 256 MiB device copy (the byte-counter "
                         "calibration kernel, copy_bytes) then one J+K step; prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the J/K path runs on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    import lightspeed_amd as ls
    from lightspeed_amd import workloads as W

    mol, bname, nocc = W.config(args.workload)
    bas = ls.Basis.from_gbs_file(mol, bname)
    D = W.random_density(bas.nao, nocc, 1)
    thresp = 1.0e-12
    t0 = time.perf_counter()
    pl = ls.PairList.build_schwarz(bas, bas, True, 1.0e-12)
    t_pl = time.perf_counter() - t0
    log("workload %s: nao=%d nprim=%d npair=%d (pair list %.2f s)" % (args.workload, bas.nao, bas.nprim, pl.npair, t_pl))
    res = ls.ResourceList.build()
    if res.ngpu == 0:
        raise RuntimeError("bench.py: no GPU visible")
    ls.hip_init(0)
    ew = ls.Ewald([1.0], [-1.0])
    Dc = np.array(ls.PureTransform.pureToCart2(bas, bas, ls.Tensor.array(D)).np) if bas.has_pure else D
    n = bas.ncart
    dD = ls.DeviceArray([n, n])
    dD.upload(Dc)
    dJ = ls.DeviceArray([n, n])
    dK = ls.DeviceArray([n, n])

    if args.pmc_pass:
        nb = 256 * 1024 * 1024 // 8
        src = ls.DeviceArray([nb])
        dst = ls.DeviceArray([nb])
        dst.copy_from(src)
        log("pmc pass: copy_bytes of %d bytes done" % (nb * 8))
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thresp, dJ.ptr)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thresp, dK.ptr)
        ls.synchronize()
        log("pmc pass: one J+K step done")
        return 0

    # quartet counts: reference convention (metric denominator) and per-launch evaluated counts
    NJ, NK = ls.count_quartets(pl, pl, dD.ptr, dD.ptr, True, thresp)
    NJ = NJ.astype(np.int64)
    NK = NK.astype(np.int64)
    ls.set_counting(True)
    ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thresp, dJ.ptr)
    evJ = ls.last_counts().astype(np.int64)
    ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thresp, dK.ptr)
    evK = ls.last_counts().astype(np.int64)
    ls.set_counting(False)
    log("quartets (reference convention): N_J=%d N_K=%d" % (NJ.sum(), NK.sum()))

    for w in range(args.warmup):
        dJ.zero()
        dK.zero()
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thresp, dJ.ptr)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thresp, dK.ptr)
        log("warmup %d done" % w)

    ls.set_profiling(True)
    launch_ms = {}
    tJ = tK = 0.0
    ls.synchronize()
    t_start = time.perf_counter()
    for s in range(args.steps):
        dJ.zero()
        dK.zero()
        a = time.perf_counter()
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thresp, dJ.ptr)
        b = time.perf_counter()
        for (i, ms) in ls.last_launches():
            launch_ms.setdefault(("J", i), []).append(ms)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thresp, dK.ptr)
        c = time.perf_counter()
        for (i, ms) in ls.last_launches():
            launch_ms.setdefault(("K", i), []).append(ms)
        tJ += b - a
        tK += c - b
    ls.synchronize()
    t_total = time.perf_counter() - t_start
    ls.set_profiling(False)
    ms_per_step = 1000.0 * t_total / args.steps
    log("timed %d steps: %.3f ms/step (J %.3f ms, K %.3f ms)" % (args.steps, ms_per_step, 1000 * tJ / args.steps,
                                                               1000 * tK / args.steps))

    # algorithmic FLOPs per launch
    def launch_flops(kind, idx):
        if kind == "J":
            L12, L34 = idx // 5, idx % 5
            f = 0.0
            for c in range(625):
                L1, L2, L3, L4 = cls_of(c)
                if L1 + L2 == L12 and L3 + L4 == L34 and NJ[c]:
                    f += NJ[c] * flops_per_quartet(L1, L2, L3, L4)
            return f
        L1, L2, L3, L4 = cls_of(idx - 256)
        return evK[idx] * flops_per_quartet(L1, L2, L3, L4)

    means = {k: float(np.mean(v)) for k, v in launch_ms.items()}
    for k in sorted(means, key=lambda k: -means[k])[:24]:
        fl = launch_flops(*k)
        name = ("J<%d,%d>" % (k[1] // 5, k[1] % 5)) if k[0] == "J" else ("K<%d,%d,%d,%d>" % cls_of(k[1] - 256))
        log("  launch %-12s %9.3f ms  %8.3f TFLOP/s  (%.3g quartets)" % (
            name, means[k], fl / (means[k] * 1e-3) / 1e12 if means[k] > 0 else 0.0,
            (sum(NJ[c] for c in range(625) if sum(cls_of(c)[:2]) == k[1] // 5 and sum(cls_of(c)[2:]) == k[1] % 5)
             if k[0] == "J" else evK[k[1]])))
    dom = max(means, key=lambda k: means[k])
    dom_fl = launch_flops(*dom)
    achieved = dom_fl / (means[dom] * 1e-3)
    total_fl = sum(NJ[c] * flops_per_quartet(*cls_of(c)) + NK[c] * flops_per_quartet(*cls_of(c)) for c in range(625))
    quartets = int(NJ.sum() + NK.sum())
    total_fl = float(total_fl)
    dom_fl = float(dom_fl)
    dom_name = ("j_quartets<%d,%d>" % (dom[1] // 5, dom[1] % 5)) if dom[0] == "J" else \
        ("k_quartets<%d,%d,%d,%d>" % cls_of(dom[1] - 256))
    roofline = {"bound": "mfma", "pipe": "fp64 (vector ALU; the MI355X FP64 peak)", "kernel": dom_name,
                "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK, "traffic": None,
                "launch_ms": means[dom], "flops_per_launch": dom_fl,
                "jk_frac": (total_fl / FP64_PEAK) / (t_total / args.steps)}

    out = {"metric": METRIC, "value": quartets / (t_total / args.steps), "unit": "ERI quartets/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "basis": bname, "nao": bas.nao, "nprim": bas.nprim,
                      "npair": pl.npair, "threpq": 1.0e-12, "thresp": thresp, "route": "fp64",
                      "J_ms": 1000 * tJ / args.steps, "K_ms": 1000 * tK / args.steps,
                      "N_J": int(NJ.sum()), "N_K": int(NK.sum()), "pairlist_s": t_pl},
           "roofline": roofline}
    if not args.no_cpu_baseline:
        log("cpu baseline (sample stride %d)" % args.cpu_stride)
        out["cpu_baseline"] = cpu_baseline(bas, D, args.cpu_stride, thresp)
        log("cpu baseline: %s" % out["cpu_baseline"]["sample"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
