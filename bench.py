#!/usr/bin/env python3
"""J+K build benchmark on one MI355X.

One step = one direct Coulomb build plus one direct exchange build (IntBox::coulomb +
IntBox::exchange, symmetric Schwarz pair list, D24symm) over an HBM-resident density, on the
workload the target is quoted on: (H2O)64 random cluster (20 A cube, seed 64), 6-31G** (1600
functions), Schwarz cutoff 1e-12, thresp = 1e-12, D = C C^T from a seeded random orthonormal C
(nocc = 320).

Two routes are timed, each over `--steps` steps after `--warmup`:
  * fp64:  thredp = thresp (every quartet FP64);
  * mixed: thredp = `--thredp` (1e-6): quartets with est >= thredp in FP64, the rest in FP32,
    accumulation FP64 (the reference's documented mixed precision, intbox.hpp:736-742).
The headline `value` is the faster route among those that meet the accuracy bar against the FP64
route on the same input (max|dJ|, max|dK| <= 1e-7, |dE2| <= 1e-6 with E2 = 2 tr(DJ) - tr(DK));
both routes' times and the errors are in `config`.

value = ERI quartets/s = (N_J + N_K) / t_JK with N counted under the reference's screening
conventions (the plain FP64 route's quartets, for both routes), t_JK the wall time of one step.
roofline: the dominant quartet kernel of the headline route (largest mean launch time, HIP
events around each launch on the stream it runs on, measured in `--attribution-steps` extra
steps after the timed region with the class launches serialised, so each duration is the
kernel's own): algorithmic FLOPs per launch = sum over its classes of N(class) x F(class),
F = 40 + nrys (30 + 9 G + 3 Pi) + 2 Pi (Pi = N1 N2 N3 N4, G = (L1+L2+1)(L3+L4+1),
nrys = (L1+L2+L3+L4)/2 + 1), against the peak of the launch's mix: its FP64 FLOPs at the FP64
peak 78.6 TFLOP/s and its FP32 FLOPs at the FP32 peak 157.3 TFLOP/s (on the mixed route one
launch per class evaluates both parts).  For a J launch the FP32 share of its classes' quartets
is the launch's evaluated FP32 fraction.
jk_frac = [sum_FP64 N F / 78.6e12 + sum_FP32 N F / 157.3e12] / t_JK.
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
FP32_PEAK = 157.3e12


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
    ap.add_argument("--thredp", type=float, default=1.0e-6, help="FP64 threshold of the mixed route")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--attribution-steps", type=int, default=2,
                    help="extra steps with serialised class launches for the per-kernel roofline")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="short run for hardware-counter passes: one 256 MiB device copy (the byte-counter "
                         "calibration kernel, copy_bytes) then one J+K step at --thredp; prints no JSON line")
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
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, args.thredp, dJ.ptr)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, args.thredp, dK.ptr)
        ls.synchronize()
        log("pmc pass: one J+K step (thredp %g) done" % args.thredp)
        return 0

    # quartet counts: reference convention (metric denominator, the plain FP64 route's quartets)
    NJ, NK = ls.count_quartets(pl, pl, dD.ptr, dD.ptr, True, thresp)
    NJ = NJ.astype(np.int64)
    NK = NK.astype(np.int64)
    quartets = int(NJ.sum() + NK.sum())
    log("quartets (reference convention): N_J=%d N_K=%d" % (NJ.sum(), NK.sum()))

    def step(thredp):
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thredp, dJ.ptr)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thredp, dK.ptr)

    # evaluated counts per route (the mixed route's FP32 share) and results for the accuracy check
    routes = {"fp64": thresp, "mixed": args.thredp}
    res_JK, cnt = {}, {}
    ls.set_counting(True)
    for name, thredp in routes.items():
        dJ.zero()
        dK.zero()
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, thredp, dJ.ptr)
        cj, cj32 = ls.last_counts().astype(np.int64), ls.last_counts_fp32().astype(np.int64)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, thredp, dK.ptr)
        ck, ck32 = ls.last_counts().astype(np.int64), ls.last_counts_fp32().astype(np.int64)
        res_JK[name] = (dJ.download(), dK.download())
        cnt[name] = (cj, cj32, ck, ck32)
    ls.set_counting(False)
    J64, K64 = res_JK["fp64"]
    Jmx, Kmx = res_JK["mixed"]
    dJmax = float(np.abs(Jmx - J64).max())
    dKmax = float(np.abs(Kmx - K64).max())
    # closed-shell two-electron energy with D = C C^T: E2 = 2 tr(D J) - tr(D K)
    dE = float(abs(2.0 * np.sum(Dc * (Jmx - J64)) - np.sum(Dc * (Kmx - K64))))
    log("mixed vs fp64: max|dJ|=%.3e max|dK|=%.3e |dE2|=%.3e" % (dJmax, dKmax, dE))
    # share of the mixed route's evaluated quartets that ran in FP32
    cjm, cjm32, ckm, ckm32 = cnt["mixed"]
    mixed_fp32 = float((cjm32.sum() + ckm32.sum()) / max(1, cjm.sum() + ckm.sum()))
    log("mixed route: FP32 share of evaluated quartets %.3f (J %.3f, K %.3f)" % (
        mixed_fp32, cjm32.sum() / max(1, cjm.sum()), ckm32.sum() / max(1, ckm.sum())))

    for w in range(args.warmup):
        for name, thredp in routes.items():
            step(thredp)
        log("warmup %d done" % w)

    # timed regions: the library's default launch concurrency (class launches over 4 streams)
    timing = {}
    for name, thredp in routes.items():
        ls.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dJ.zero()
            dK.zero()
            step(thredp)
        ls.synchronize()
        timing[name] = (time.perf_counter() - t0) / args.steps
        log("%s: %d steps, %.3f ms/step" % (name, args.steps, 1000 * timing[name]))
    # headline route: the faster of the two among those that meet the accuracy bar (fp64 always does)
    mixed_ok = dJmax <= 1e-7 and dKmax <= 1e-7 and dE <= 1e-6
    head = "mixed" if (mixed_ok and timing["mixed"] < timing["fp64"]) else "fp64"
    t_step = timing[head]

    # kernel attribution: the headline route with the class launches serialised on one stream, so
    # that each launch's HIP-event duration is its own (under concurrency launches share the device)
    ls.set_streams(1)
    ls.set_profiling(True)
    launch_ms = {}
    for _ in range(args.attribution_steps):
        dJ.zero()
        dK.zero()
        ls.IntBox.coulomb_device(ew, pl, pl, dD.ptr, thresp, routes[head], dJ.ptr)
        for (i, ms) in ls.last_launches():
            launch_ms.setdefault(("J", i), []).append(ms)
        ls.IntBox.exchange_device(ew, pl, pl, dD.ptr, True, thresp, routes[head], dK.ptr)
        for (i, ms) in ls.last_launches():
            launch_ms.setdefault(("K", i), []).append(ms)
    ls.set_profiling(False)
    ls.set_streams(4)

    # algorithmic FLOPs (frozen convention) per class, split by route
    def cls_flops(c):
        return float(flops_per_quartet(*cls_of(c)))

    cj, cj32, ck, ck32 = cnt[head]

    def launch_flops(kind, lid):
        """(FP64 FLOPs, FP32 FLOPs) of one launch."""
        idx = lid
        # one launch per class evaluates both parts of the mixed route
        if kind == "J":
            # a symmetric J launch (bit 24) evaluates two group pairs (the second id, + 1, in bits
            # 12..23)
            f64 = f32s = 0.0
            for i in (idx & 0xfff, ((idx >> 12) & 0xfff) - 1):
                if i < 0:
                    continue
                L12, L34 = i // 5, i % 5
                full = sum(NJ[c] * cls_flops(c) for c in range(625)
                           if sum(cls_of(c)[:2]) == L12 and sum(cls_of(c)[2:]) == L34)
                f32 = cj32[i] / cj[i] if cj[i] else 0.0
                f64 += full * (1.0 - f32)
                f32s += full * f32
            return f64, f32s
        c = idx - 256
        return (ck[idx] - ck32[idx]) * cls_flops(c), ck32[idx] * cls_flops(c)

    def launch_name(kind, lid):
        idx = lid
        # template arguments as the profiler prints them: route 0 plain FP64, 4 mixed (FP64 and FP32
        # parts in one launch)
        if kind == "J":
            if (idx >> 24) & 1:
                return "j_quartets_sym<%d, %d>" % ((idx & 0xfff) // 5, (idx & 0xfff) % 5)
            return "j_quartets<%d, %d, %d>" % (idx // 5, idx % 5, 4 if head == "mixed" else 0)
        return "k_ket<%d, %d, %d, %d, %d>" % (cls_of(idx - 256) + (4 if head == "mixed" else 0,))

    means = {k: float(np.mean(v)) for k, v in launch_ms.items()}
    log("serialised quartet kernels per step: J %.3f ms, K %.3f ms" % (
        sum(v for k, v in means.items() if k[0] == "J"), sum(v for k, v in means.items() if k[0] == "K")))
    for k in sorted(means, key=lambda k: -means[k]):
        fl = sum(launch_flops(*k))
        log("  launch %-36s %9.3f ms  %8.3f TFLOP/s" % (launch_name(*k), means[k], fl / (means[k] * 1e-3) / 1e12))
    dom = max(means, key=lambda k: means[k])
    dom64, dom32 = launch_flops(*dom)
    dom_fl = dom64 + dom32
    # peak of the launch's mix: FP64 FLOPs at the FP64 peak, FP32 FLOPs at the FP32 peak
    peak = dom_fl / (dom64 / FP64_PEAK + dom32 / FP32_PEAK) if dom_fl > 0 else FP64_PEAK
    achieved = dom_fl / (means[dom] * 1e-3)
    dom_lo = dom32 > 0.0
    # whole-step fraction: sum_FP64 N F / 78.6e12 + sum_FP32 N F / 157.3e12, over the step time
    n32 = {c: 0.0 for c in range(625)}
    for c in range(625):
        n32[c] = float(ck32[256 + c])
    for idx in range(25):
        if cj[idx]:
            f32 = cj32[idx] / cj[idx]
            for c in range(625):
                L1, L2, L3, L4 = cls_of(c)
                if (L1 + L2) * 5 + (L3 + L4) == idx:
                    n32[c] += f32 * NJ[c]
    t_ideal = 0.0
    for c in range(625):
        F = cls_flops(c)
        tot = float(NJ[c] + NK[c])
        t_ideal += (tot - n32[c]) * F / FP64_PEAK + n32[c] * F / FP32_PEAK
    roofline = {"bound": "mfma", "pipe": "%s vector ALU (MI355X peak)" % ("fp32" if dom_lo else "fp64"),
                "kernel": launch_name(*dom), "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "launch_ms": means[dom], "flops_per_launch": float(dom_fl),
                "launch_timing": "HIP events around each launch on its stream, class launches serialised "
                                 "(%d steps after the timed region)" % args.attribution_steps,
                "jk_frac": t_ideal / t_step}
    fp32_share = float(sum(n32.values()) / quartets) if head == "mixed" else 0.0
    # HBM bytes per launch of the dominant kernel, from the committed counter summary of this
    # workload (two rocprofv3 passes over `bench.py --pmc-pass`, calibrated on copy_bytes)
    traffic_md = os.path.join(ROOT, "profiles", "jk_%s_hbm_traffic.md" % args.workload)
    if os.path.exists(traffic_md):
        for line in open(traffic_md):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0] == roofline["kernel"]:
                roofline["traffic"] = float(cells[4]) * 1e6
                roofline["traffic_source"] = os.path.relpath(traffic_md, ROOT)

    out = {"metric": METRIC, "value": quartets / t_step, "unit": "ERI quartets/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000 * t_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32" if head == "mixed" else "f64",
           "data": "synthetic",
           "config": {"workload": args.workload, "basis": bname, "nao": bas.nao, "nprim": bas.nprim,
                      "npair": pl.npair, "threpq": 1.0e-12, "thresp": thresp, "route": head,
                      "thredp": routes[head], "N_J": int(NJ.sum()), "N_K": int(NK.sum()),
                      "fp32_quartet_share": fp32_share,
                      "fp64_ms_per_step": 1000 * timing["fp64"], "fp64_quartets_per_s": quartets / timing["fp64"],
                      "mixed_ms_per_step": 1000 * timing["mixed"],
                      "mixed_vs_fp64_speedup": timing["fp64"] / timing["mixed"],
                      "mixed_fp32_quartet_share": mixed_fp32,
                      "mixed_meets_accuracy": bool(mixed_ok), "mixed_max_abs_dJ": dJmax, "mixed_max_abs_dK": dKmax, "mixed_abs_dE2": dE,
                      "pairlist_s": t_pl},
           "roofline": roofline}
    if not args.no_cpu_baseline:
        log("cpu baseline (sample stride %d)" % args.cpu_stride)
        out["cpu_baseline"] = cpu_baseline(bas, D, args.cpu_stride, thresp)
        log("cpu baseline: %s" % out["cpu_baseline"]["sample"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
