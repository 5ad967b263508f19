#!/usr/bin/env python3
"""J+K build benchmark on one MI355X.

One step = one direct Coulomb build plus one direct exchange build (IntBox::coulomb +
IntBox::exchange, symmetric Schwarz pair list, D24symm) over an HBM-resident density, on the
workload the target is quoted on: (H2O)64 random cluster (20 A cube, seed 64), 6-31G** (1600
functions), Schwarz cutoff 1e-12, thresp = 1e-12, D = C C^T from a seeded random orthonormal C
(nocc = 320).

Routes, each timed over `--steps` steps after `--warmup`:
  * fp64:  thredp = thresp (every quartet FP64);
  * mixed: thredp = `--thredp` (1e-6): quartets with est >= thredp in FP64, the rest in FP32,
    accumulation FP64 (the reference's documented mixed precision, intbox.hpp:736-742);
  * learned, learned_mixed: J on the FP64 route, K through the trained screen
    (lightspeed_amd/data/screen_h2o_631gss.npz, or --screen) at thredp = thresp / `--thredp`.
Every route's J and K are compared with the FP64 route (whole matrices, and the two-electron energy
E2 = 2 tr(DJ) - tr(DK)) and, on the CPU-baseline leg's bra-pair sample, with the FP64 oracle.  The
headline `value` is the mixed route when it meets the accuracy bar (max|dJ|, max|dK| <= 1e-7,
|dE2| <= 1e-6; the metric compares it with the FP64 route), else the FP64 route; `roofline` is the
headline route's and `roofline_fp64_route` the FP64 route's when they differ.  Before the
timed region the autotuner picks the J ket chunks and K unit splits on this workload, then for the
mixed routes the faster of the mixed and the FP64 kernel per launch (the default parameters' FP64
time is reported beside it; --no-autotune skips it).

value = ERI quartets/s = (N_J + N_K) / t_JK with N counted under the reference's screening
conventions (lsjk_count_quartets: the plain FP64 route's quartets, for every route), t_JK the wall
time of one step.
roofline: the dominant quartet kernel of the headline route (largest mean launch time, HIP events
around each launch on the stream it runs on, measured in `--attribution-steps` extra steps after
the timed region with the class launches serialised, so each duration is the kernel's own):
algorithmic FLOPs per launch = sum over its classes of N(class) x F(class) with N from
lsjk_count_quartets (a K launch of a class with a distinct mirror class evaluates both, N(class) +
N(mirror)), F = 40 + nrys (30 + 9 G + 3 Pi) + 2 Pi (Pi = N1 N2 N3 N4, G = (L1+L2+1)(L3+L4+1),
nrys = (L1+L2+L3+L4)/2 + 1), against the peak of the launch's mix: its FP64 FLOPs at the FP64 peak
78.6 TFLOP/s and its FP32 FLOPs (the launch's evaluated FP32 share) at the FP32 peak 157.3 TFLOP/s.
jk_frac = [sum_FP64 N F / 78.6e12 + sum_FP32 N F / 157.3e12] / t_JK (algorithmic counts);
jk_exec_frac = the same over the quartets the kernels evaluated (their own counters).
`--table PATH` writes the headline route's per-launch table (N, F, evaluated, ms, TFLOP/s, fraction,
lane use = the share of the wave evaluating integrals in its busy steps), and PATH_fp64 the FP64
route's when the headline is another route.
cpu_baseline: the reference's own CPU IntBox::coulombCPU + IntBox::exchangeCPU, compiled from
its sources into oracle/_ref/ref_jk, timed on a bounded sample of the same workload (the bra list
keeps every `--cpu-stride`-th pair of each class, the ket list is whole), all threads of the host.
The sample's quartet count comes from the CPU FP64 oracle run on the same sample (its counts
equal the reference's evaluated quartets); where the reference build is absent the oracle's own
time is reported (kind "port").
Other workloads (config.other_workloads; --quick skips them): C20H42/cc-pVDZ and (H2O)256/6-31G*
(autotuned as above; FP64 and mixed ms per step, quartets/s, jk_frac, reference CPU rate on a
bra-pair sample), and the
batched workload, 256 seeded (H2O)n, n = 1..16, 6-31G*, all J and K builds in one graph replay
(molecules/s, FP64 with default and autotuned parameters, and mixed).

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload h2o64_631gss] [--quick]
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


def mirror(idx):
    """Counter index of the mirror class (L3 L4 | L1 L2) of K counter index idx (256 + class)."""
    L1, L2, L3, L4 = cls_of(idx - 256)
    return 256 + ((L3 * 5 + L4) * 5 + L1) * 5 + L2


def cpu_baseline(bas, D, stride, thresp, accuracy=False):
    """Time the reference build (or the oracle) on a bra-pair sample; the oracle supplies the count.
    accuracy: also return the oracle's J and K (D24symm off, so the sample's K is the plain sum) on
    the same sample, for the routes' errors against the FP64 oracle."""
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
    Jo = O.coulomb(ew, sub, opl, D, thresp, counts=cj)
    t1 = time.perf_counter()
    O.exchange(ew, sub, opl, D, True, thresp, counts=ck)
    t2 = time.perf_counter()
    Ko = O.exchange(ew, sub, opl, D, False, thresp) if accuracy else None
    n = int(cj.sum() + ck.sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    port = "CPU FP64 oracle (OpenMP, %d threads): J %.2f s + K %.2f s" % (threads, t1 - t0, t2 - t1)
    sample = "every %d-th bra pair of each class of the same workload, whole ket list: %d J + %d K quartets" % (
        stride, int(cj.sum()), int(ck.sum()))
    if refjk.available():
        _, _, (tj, tk, _) = refjk.run(bas, D, 1.0e-12, thresp, symmetric=True, D24symm=True, stride=stride,
                                      nthreads=threads)
        out = {"value": n / (tj + tk), "unit": "ERI quartets/s", "cores": threads, "kind": "reference",
               "sample": "reference IntBox::coulombCPU + exchangeCPU built from its sources (OpenMP, %d "
                         "threads) on %s in J %.2f s + K %.2f s; quartets counted by the oracle on the same "
                         "sample (%s)" % (threads, sample, tj, tk, port)}
    else:
        out = {"value": n / (t2 - t0), "unit": "ERI quartets/s", "cores": threads, "kind": "port",
               "sample": "%s on %s" % (port, sample)}
    return out, (Jo, Ko)


class Workload:
    """A single-molecule workload on the device: basis, density, pair list, HBM buffers, quartet
    counts under the reference's conventions (the metric's denominator)."""

    def __init__(self, ls, W, name, thresp):
        self.ls, self.name, self.thresp = ls, name, thresp
        mol, self.bname, nocc = W.config(name)
        self.bas = ls.Basis.from_gbs_file(mol, self.bname)
        self.D = W.random_density(self.bas.nao, nocc, 1)
        t0 = time.perf_counter()
        self.pl = ls.PairList.build_schwarz(self.bas, self.bas, True, 1.0e-12)
        self.t_pl = time.perf_counter() - t0
        self.ew = ls.Ewald([1.0], [-1.0])
        b = self.bas
        self.Dc = np.array(ls.PureTransform.pureToCart2(b, b, ls.Tensor.array(self.D)).np) if b.has_pure else self.D
        n = b.ncart
        self.dD = ls.DeviceArray([n, n])
        self.dD.upload(self.Dc)
        self.dJ = ls.DeviceArray([n, n])
        self.dK = ls.DeviceArray([n, n])
        NJ, NK = ls.count_quartets(self.pl, self.pl, self.dD.ptr, self.dD.ptr, True, thresp)
        self.NJ, self.NK = NJ.astype(np.int64), NK.astype(np.int64)
        self.quartets = int(self.NJ.sum() + self.NK.sum())
        log("workload %s: nao=%d nprim=%d npair=%d (pair list %.2f s); quartets (reference convention): "
            "N_J=%d N_K=%d" % (name, b.nao, b.nprim, self.pl.npair, self.t_pl, self.NJ.sum(), self.NK.sum()))

    def step(self, route, pl12=None, D24symm=True, dJ=None, dK=None):
        """One J + K build on the route (thredp, screen): J on the plain route when the route has a
        screen (the screen certifies exchange classes only)."""
        ls = self.ls
        thredp, screen = route
        pl12 = pl12 or self.pl
        dJ, dK = dJ or self.dJ, dK or self.dK
        ls.IntBox.coulomb_device(self.ew, pl12, self.pl, self.dD.ptr, self.thresp,
                                 self.thresp if screen else thredp, dJ.ptr)
        if screen:
            screen.exchange_device(self.ew, pl12, self.pl, self.dD.ptr, D24symm, self.thresp, thredp, dK.ptr)
        else:
            ls.IntBox.exchange_device(self.ew, pl12, self.pl, self.dD.ptr, D24symm, self.thresp, thredp, dK.ptr)

    def time_routes(self, routes, steps, warmup):
        ls = self.ls
        for w in range(warmup):
            for r in routes.values():
                self.step(r)
        timing = {}
        for name, r in routes.items():
            ls.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                self.dJ.zero()
                self.dK.zero()
                self.step(r)
            ls.synchronize()
            timing[name] = (time.perf_counter() - t0) / steps
            log("%s %s: %d steps, %.3f ms/step" % (self.name, name, steps, 1000 * timing[name]))
        return timing

    def ideal_time(self, n32=None):
        """sum_FP64 N F / 78.6e12 + sum_FP32 N F / 157.3e12 over the reference-convention counts
        (n32: per class the part evaluated in FP32)."""
        t = 0.0
        for c in range(625):
            F = float(flops_per_quartet(*cls_of(c)))
            tot = float(self.NJ[c] + self.NK[c])
            f32 = n32[c] if n32 is not None else 0.0
            t += (tot - f32) * F / FP64_PEAK + f32 * F / FP32_PEAK
        return t


def measure_extra(ls, W, name, steps, cpu_stride, thredp, tune=True):
    """FP64 and mixed routes of another single-molecule workload: ms per step, quartets/s, the
    whole-step roofline fraction, and the reference CPU rate on a bra-pair sample.  With tune, the
    autotuner runs first (FP64 launch shapes, then the mixed route's kernel per launch)."""
    wl = Workload(ls, W, name, 1.0e-12)
    routes = {"fp64": (wl.thresp, None), "mixed": (thredp, None)}
    t_tune = None
    if tune:
        t0 = time.perf_counter()
        ls.autotune(wl.pl, wl.ew, wl.dD.ptr, wl.thresp, wl.thresp)
        ls.autotune(wl.pl, wl.ew, wl.dD.ptr, wl.thresp, thredp)
        t_tune = time.perf_counter() - t0
        log("%s autotune: %.1f s" % (name, t_tune))
    res = {}
    for r in routes:
        wl.dJ.zero()
        wl.dK.zero()
        wl.step(routes[r])
        res[r] = (wl.dJ.download(), wl.dK.download())
    timing = wl.time_routes(routes, steps, 1)
    dJ = float(np.abs(res["mixed"][0] - res["fp64"][0]).max())
    dK = float(np.abs(res["mixed"][1] - res["fp64"][1]).max())
    out = {"basis": wl.bname, "nao": wl.bas.nao, "ncart": wl.bas.ncart, "nprim": wl.bas.nprim, "npair": wl.pl.npair,
           "N_J": int(wl.NJ.sum()), "N_K": int(wl.NK.sum()), "steps": steps,
           "fp64_ms_per_step": 1000 * timing["fp64"], "fp64_quartets_per_s": wl.quartets / timing["fp64"],
           "fp64_jk_frac": wl.ideal_time() / timing["fp64"],
           "mixed_ms_per_step": 1000 * timing["mixed"], "mixed_quartets_per_s": wl.quartets / timing["mixed"],
           "mixed_max_abs_dJ_vs_fp64": dJ, "mixed_max_abs_dK_vs_fp64": dK, "autotune_s": t_tune}
    if cpu_stride:
        log("%s: cpu baseline (sample stride %d)" % (name, cpu_stride))
        cb, _ = cpu_baseline(wl.bas, wl.D, cpu_stride, wl.thresp)
        out["cpu_baseline"] = cb
        log("%s: cpu baseline %.3g quartets/s (%s)" % (name, cb["value"], cb["kind"]))
    return out


def measure_batch(ls, W, nmol, steps, thredp, tune=True):
    """The batched workload: nmol seeded (H2O)_n, n uniform in 1..16, 6-31G*, nocc = 5n, all J and K
    builds in one graph replay (FP64 and mixed routes); molecules/s."""
    t0 = time.perf_counter()
    res = ls.ResourceList.build()
    mols = []
    for mol, nocc, dseed in W.water_batch(nmol):
        bas = ls.Basis.from_gbs_file(mol, "6-31gs")
        mols.append((bas, W.random_density(bas.nao, nocc, dseed), ls.PairList.build_schwarz(bas, bas, True, 1e-12)))
    batch = ls.JKBatch.build(res, [pl for _, _, pl in mols])
    t_build = time.perf_counter() - t0
    buf = np.zeros(batch.nelem)
    for off, (bas, D, _) in zip(batch.offsets, mols):
        n = bas.ncart
        for i in range(n):
            buf[off + i * batch.ld: off + i * batch.ld + n] = D[i]
    dD = ls.DeviceArray([batch.nelem])
    dD.upload(buf)
    dJ = ls.DeviceArray([batch.nelem])
    dK = ls.DeviceArray([batch.nelem])
    ew = ls.Ewald([1.0], [-1.0])
    log("batch: %d molecules, %d functions, largest %d (host build %.1f s)" % (
        nmol, sum(b.ncart for b, _, _ in mols), max(b.ncart for b, _, _ in mols), t_build))
    out = {"molecules": nmol, "basis": "6-31gs", "functions": int(sum(b.ncart for b, _, _ in mols)),
           "largest": int(max(b.ncart for b, _, _ in mols)), "host_build_s": t_build, "steps": steps}
    passes = [("fp64_default", 1e-12), ("fp64", 1e-12), ("mixed", thredp)] if tune else [("fp64", 1e-12),
                                                                                          ("mixed", thredp)]
    for name, td in passes:
        if name == "fp64" and tune:
            t0 = time.perf_counter()
            table = batch.autotune(ew, dD.ptr, 1e-12, 1e-12)
            batch.autotune(ew, dD.ptr, 1e-12, thredp)  # the mixed route's kernel per launch
            out["autotune_s"] = time.perf_counter() - t0
            out["autotune_changed"] = sum(1 for k, i, v, a, b in table if (k < 2 and v != 1.0) or (k == 2 and v != 12288) or (k == 4 and b < a))
            log("batch autotune: %.1f s, %d of %d entries changed" % (out["autotune_s"], out["autotune_changed"], len(table)))
        t0 = time.perf_counter()
        batch.jk_device(ew, dD.ptr, 1e-12, td, dJ.ptr, dK.ptr)  # eager run + graph capture
        out[name + "_first_call_s"] = time.perf_counter() - t0
        ls.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dJ.zero()
            dK.zero()
            batch.jk_device(ew, dD.ptr, 1e-12, td, dJ.ptr, dK.ptr)
        ls.synchronize()
        t = (time.perf_counter() - t0) / steps
        out[name + "_ms_per_step"] = 1000 * t
        out[name + "_molecules_per_s"] = nmol / t
        log("batch %s: %.1f ms per graph replay, %.1f molecules/s" % (name, 1000 * t, nmol / t))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o64_631gss")
    ap.add_argument("--cpu-stride", type=int, default=32)
    ap.add_argument("--thredp", type=float, default=1.0e-6, help="FP64 threshold of the mixed routes")
    ap.add_argument("--screen", default=None, help="trained screen (.npz; default: the package's model)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline workload only (no other configurations)")
    ap.add_argument("--no-autotune", action="store_true", help="time the default launch parameters only")
    ap.add_argument("--table", default=None, help="write the per-launch table (markdown) here")
    ap.add_argument("--attribution-steps", type=int, default=2,
                    help="extra steps with serialised class launches for the per-kernel roofline")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="short run for hardware-counter passes: one 256 MiB device copy (the byte-counter "
                         "calibration kernel, copy_bytes) then one J+K step at --thredp; prints no JSON line")
    ap.add_argument("--profile-marker", action="store_true",
                    help="launch one 8-byte copy_bytes right before the headline's timed steps, as a cut point "
                         "for a kernel trace of the run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the J/K path runs on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2

    import lightspeed_amd as ls
    from lightspeed_amd import workloads as W
    from lightspeed_amd import screen as SC

    res = ls.ResourceList.build()
    if res.ngpu == 0:
        raise RuntimeError("bench.py: no GPU visible")
    ls.hip_init(0)
    thresp = 1.0e-12
    wl = Workload(ls, W, args.workload, thresp)
    bas, pl, Dc = wl.bas, wl.pl, wl.Dc
    NJ, NK, quartets = wl.NJ, wl.NK, wl.quartets

    if args.pmc_pass:
        nb = 256 * 1024 * 1024 // 8
        src = ls.DeviceArray([nb])
        dst = ls.DeviceArray([nb])
        dst.copy_from(src)
        log("pmc pass: copy_bytes of %d bytes done" % (nb * 8))
        wl.step((args.thredp, None))
        ls.synchronize()
        log("pmc pass: one J+K step (thredp %g) done" % args.thredp)
        return 0

    # the learned screen (trained on seeded (H2O)8..32 clusters by python -m lightspeed_amd.screen)
    screen_path = args.screen or SC.DEFAULT_MODEL
    screen = SC.Screen.load(screen_path) if os.path.exists(screen_path) else None
    routes = {"fp64": (thresp, None), "mixed": (args.thredp, None)}
    if screen is not None:
        routes["learned"] = (thresp, screen)
        routes["learned_mixed"] = (args.thredp, screen)
    else:
        log("no trained screen at %s: learned routes not measured" % screen_path)

    # the autotuner (J ket chunks per group pair, K ket-group rows and unit splits, timed on this
    # list and density on the FP64 route, then the mixed route's kernel per launch), before the
    # counted calls and accuracy checks, so that these describe the kernels the timed steps run;
    # the default parameters' FP64 time is measured first
    tune_info = None
    if not args.no_autotune:
        t_def = wl.time_routes({"fp64_default": routes["fp64"]}, args.steps, args.warmup)["fp64_default"]
        t0 = time.perf_counter()
        table = ls.autotune(pl, wl.ew, wl.dD.ptr, thresp, thresp)
        # the mixed route: per launch, its mixed kernel or the FP64 kernel, whichever is faster
        table_mixed = ls.autotune(pl, wl.ew, wl.dD.ptr, thresp, args.thredp)
        n_plain = sum(1 for k, i, v, _, _ in table_mixed if k == 3 and v == 1.0)
        log("autotune, mixed route: %d of %d launches on the FP64 kernel" % (
            n_plain, sum(1 for k, i, v, _, _ in table_mixed if k == 3)))
        t_tune = time.perf_counter() - t0
        changed = [(k, i, v) for k, i, v, a, b in table if (k < 2 and v != 1.0) or (k == 2 and v != 12288) or (k == 4 and b < a)]
        log("autotune: %.1f s, %d entries, %d changed" % (t_tune, len(table), len(changed)))
        tune_info = {"seconds": t_tune, "entries": len(table), "changed": len(changed),
                     "mixed_route_fp64_kernel_launches": n_plain,
                     "fp64_default_ms_per_step": 1000 * t_def,
                     "table": [[int(k), int(i), float(v), float(a), float(b)] for k, i, v, a, b in table]}
    # evaluated counts per route and results for the accuracy checks (whole matrices vs the FP64 route)
    res_JK, cnt, occ = {}, {}, {}
    ls.set_counting(True)
    for name, r in routes.items():
        wl.dJ.zero()
        wl.dK.zero()
        thredp, scr = r
        ls.IntBox.coulomb_device(wl.ew, pl, pl, wl.dD.ptr, thresp, thresp if scr else thredp, wl.dJ.ptr)
        cj, cj32 = ls.last_counts().astype(np.int64), ls.last_counts_fp32().astype(np.int64)
        wsj, wlj = (x.astype(np.int64) for x in ls.last_wave_steps())
        if scr:
            scr.exchange_device(wl.ew, pl, pl, wl.dD.ptr, True, thresp, thredp, wl.dK.ptr)
        else:
            ls.IntBox.exchange_device(wl.ew, pl, pl, wl.dD.ptr, True, thresp, thredp, wl.dK.ptr)
        ck, ck32 = ls.last_counts().astype(np.int64), ls.last_counts_fp32().astype(np.int64)
        wsk, wlk = (x.astype(np.int64) for x in ls.last_wave_steps())
        res_JK[name] = (wl.dJ.download(), wl.dK.download())
        cnt[name] = (cj, cj32, ck, ck32)
        occ[name] = (wsj, wlj, wsk, wlk)
    ls.set_counting(False)
    J64, K64 = res_JK["fp64"]
    acc = {}
    for name in routes:
        J, K = res_JK[name]
        acc[name] = {"max_abs_dJ_vs_fp64": float(np.abs(J - J64).max()),
                     "max_abs_dK_vs_fp64": float(np.abs(K - K64).max()),
                     # closed-shell two-electron energy with D = C C^T: E2 = 2 tr(D J) - tr(D K)
                     "abs_dE2_vs_fp64": float(abs(2.0 * np.sum(Dc * (J - J64)) - np.sum(Dc * (K - K64)))),
                     "evaluated_quartets": int(cnt[name][0].sum() + cnt[name][2].sum()),
                     "fp32_share": float((cnt[name][1].sum() + cnt[name][3].sum()) /
                                         max(1, cnt[name][0].sum() + cnt[name][2].sum()))}
        log("%s: evaluated %d quartets (FP32 share %.3f); vs fp64: max|dJ|=%.2e max|dK|=%.2e |dE2|=%.2e" % (
            name, acc[name]["evaluated_quartets"], acc[name]["fp32_share"], acc[name]["max_abs_dJ_vs_fp64"],
            acc[name]["max_abs_dK_vs_fp64"], acc[name]["abs_dE2_vs_fp64"]))

    # the CPU baseline leg (reference build timed on a bra-pair sample) and, on the same sample, every
    # route against the FP64 oracle
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (sample stride %d)" % args.cpu_stride)
        cpu, (Jo, Ko) = cpu_baseline(bas, wl.D, args.cpu_stride, thresp, accuracy=True)
        log("cpu baseline: %s" % cpu["sample"])
        sub = pl.subsample(args.cpu_stride)
        n = bas.ncart
        sJ, sK = ls.DeviceArray([n, n]), ls.DeviceArray([n, n])
        T = ls.PureTransform
        for name, r in routes.items():
            sJ.zero()
            sK.zero()
            wl.step(r, pl12=sub, D24symm=False, dJ=sJ, dK=sK)
            J, K = sJ.download(), sK.download()
            if bas.has_pure:
                J = np.array(T.cartToPure2(bas, bas, ls.Tensor.array(J), ls.Tensor.array(np.zeros_like(wl.D))).np)
                K = np.array(T.cartToPure2(bas, bas, ls.Tensor.array(K), ls.Tensor.array(np.zeros_like(wl.D))).np)
            acc[name]["max_abs_dJ_vs_oracle_sample"] = float(np.abs(J - Jo).max())
            acc[name]["max_abs_dK_vs_oracle_sample"] = float(np.abs(K - Ko).max())
            log("%s vs FP64 oracle (sample): max|dJ|=%.2e max|dK|=%.2e" % (
                name, acc[name]["max_abs_dJ_vs_oracle_sample"], acc[name]["max_abs_dK_vs_oracle_sample"]))

    if args.profile_marker:
        # a one-element device copy (kernel copy_bytes) right before the timed steps: a kernel trace
        # of this run is cut there, so that its per-kernel durations cover the timed and attribution
        # steps only (autotuned launch parameters), not the tuning and counting calls before them
        marker = ls.DeviceArray([1])
        marker.copy_from(ls.DeviceArray([1]))
        ls.synchronize()
    timing = wl.time_routes(routes, args.steps, args.warmup)

    # headline route: the fastest among those that meet the accuracy bar (max|dJ|, max|dK| <= 1e-7
    # against the FP64 route and, where measured, the FP64 oracle; |dE2| <= 1e-6); fp64 always does
    def meets(name):
        a = acc[name]
        ok = a["max_abs_dJ_vs_fp64"] <= 1e-7 and a["max_abs_dK_vs_fp64"] <= 1e-7 and a["abs_dE2_vs_fp64"] <= 1e-6
        if "max_abs_dK_vs_oracle_sample" in a:
            ok = ok and a["max_abs_dJ_vs_oracle_sample"] <= 1e-7 and a["max_abs_dK_vs_oracle_sample"] <= 1e-7
        return ok

    # the headline is the metric's subject, the mixed route, when it meets the accuracy bar (else
    # the FP64 route); the FP64 route's time and roofline are reported beside it
    head = "mixed" if "mixed" in routes and meets("mixed") else "fp64"
    t_step = timing[head]

    def attribute(rname, t_route, table_path):
        """Per-launch attribution and the roofline of route rname (see the module docstring)."""
        # kernel attribution: the route with the class launches serialised on one stream, so that
        # each launch's HIP-event duration is its own (under concurrency launches share the device)
        ls.set_streams(1)
        ls.set_profiling(True)
        launch_ms = {}
        for _ in range(args.attribution_steps):
            wl.dJ.zero()
            wl.dK.zero()
            thredp, scr = routes[rname]
            ls.IntBox.coulomb_device(wl.ew, pl, pl, wl.dD.ptr, thresp, thresp if scr else thredp, wl.dJ.ptr)
            for (i, ms) in ls.last_launches():
                launch_ms.setdefault(("J", i), []).append(ms)
            if scr:
                scr.exchange_device(wl.ew, pl, pl, wl.dD.ptr, True, thresp, thredp, wl.dK.ptr)
            else:
                ls.IntBox.exchange_device(wl.ew, pl, pl, wl.dD.ptr, True, thresp, thredp, wl.dK.ptr)
            for (i, ms) in ls.last_launches():
                launch_ms.setdefault(("K", i), []).append(ms)
        ls.set_profiling(False)
        ls.set_streams(4)

        def cls_flops(c):
            return float(flops_per_quartet(*cls_of(c)))

        cj, cj32, ck, ck32 = cnt[rname]
        j_mixed = routes[rname][0] > thresp and routes[rname][1] is None

        def launch_work(kind, lid):
            """(N, F-weighted FP64 FLOPs, FP32 FLOPs, evaluated quartets) of one launch: N and the FLOPs
            under the reference's conventions (lsjk_count_quartets), the FP32 part in the proportion the
            launch's kernels evaluated in FP32."""
            if kind == "J":
                # a symmetric J launch (bit 24) evaluates two group pairs (the second id, + 1, in bits 12..23)
                N = f64 = f32s = 0.0
                ev = 0
                for i in (lid & 0xfff, ((lid >> 12) & 0xfff) - 1):
                    if i < 0 or i >= 25:
                        continue
                    L12, L34 = i // 5, i % 5
                    cs = [c for c in range(625) if sum(cls_of(c)[:2]) == L12 and sum(cls_of(c)[2:]) == L34]
                    n_ = float(sum(NJ[c] for c in cs))
                    full = sum(NJ[c] * cls_flops(c) for c in cs)
                    f32 = cj32[i] / cj[i] if cj[i] else 0.0
                    N += n_
                    f64 += full * (1.0 - f32)
                    f32s += full * f32
                    ev += int(cj[i])
                return N, f64, f32s, ev
            # a K launch of a class with a distinct mirror class evaluates both (symmetric mode)
            c = lid - 256
            m = mirror(lid)
            N = float(NK[c] + (NK[m - 256] if m != lid else 0))
            # the counts sit at the index of the orientation that ran in the counting call (without
            # the autotuner the first call picks it): take both indices
            n_ev = int(ck[lid] + (ck[m] if m != lid else 0))
            n_32 = int(ck32[lid] + (ck32[m] if m != lid else 0))
            f32 = n_32 / n_ev if n_ev else 0.0
            return N, N * cls_flops(c) * (1.0 - f32), N * cls_flops(c) * f32, n_ev

        def launch_name(kind, lid):
            # template arguments as the profiler prints them: route 0 plain FP64, 4 mixed (FP64 and FP32
            # parts in one launch)
            if kind == "J":
                if (lid >> 24) & 1:
                    return "j_quartets_sym<%d, %d>" % ((lid & 0xfff) // 5, (lid & 0xfff) % 5)
                return "j_quartets<%d, %d, %d>" % (lid // 5, lid % 5, 4 if j_mixed else 0)
            m = mirror(lid)
            k_route = 4 if (routes[rname][0] > thresp and ck32[lid] + (ck32[m] if m != lid else 0) > 0) else 0
            return "k_ket<%d, %d, %d, %d, %d>" % (cls_of(lid - 256) + (k_route,))

        def lane_use(kind, lid):
            """Share of the wave's lanes evaluating an integral in the launch's busy wave steps."""
            wsj, wlj, wsk, wlk = occ[rname]
            if kind == "J":
                i = lid & 0xfff
                steps, lanes = wsj[i], wlj[i]
            else:
                m = mirror(lid)
                steps = wsk[lid] + (wsk[m] if m != lid else 0)
                lanes = wlk[lid] + (wlk[m] if m != lid else 0)
            return float(lanes) / (64.0 * steps) if steps else float("nan")

        means = {k: float(np.mean(v)) for k, v in launch_ms.items()}
        log("%s route, serialised quartet kernels per step: J %.3f ms, K %.3f ms" % (rname,
            sum(v for k, v in means.items() if k[0] == "J"), sum(v for k, v in means.items() if k[0] == "K")))
        rows = []
        for k in sorted(means, key=lambda k: -means[k]):
            N, f64, f32, ev = launch_work(*k)
            fl = f64 + f32
            pk = fl / (f64 / FP64_PEAK + f32 / FP32_PEAK) if fl > 0 else FP64_PEAK
            rows.append((launch_name(*k), N, fl / max(N, 1.0), ev, means[k], fl / (means[k] * 1e-3) / 1e12,
                         fl / (means[k] * 1e-3) / pk, lane_use(*k)))
            log("  launch %-36s %9.3f ms  %8.3f TFLOP/s" % (rows[-1][0], means[k], rows[-1][5]))
        if table_path:
            with open(table_path, "w") as f:
                f.write("| launch | N (reference convention) | FLOP/quartet | evaluated | ms | TFLOP/s | frac "
                        "| lane use |\n")
                f.write("|---|---:|---:|---:|---:|---:|---:|---:|\n")
                for r in rows:
                    f.write("| %s | %d | %.0f | %d | %.3f | %.2f | %.3f | %.3f |\n" % r)
        dom = max(means, key=lambda k: means[k])
        Nd, dom64, dom32, _ = launch_work(*dom)
        dom_fl = dom64 + dom32
        # peak of the launch's mix: FP64 FLOPs at the FP64 peak, FP32 FLOPs at the FP32 peak
        peak = dom_fl / (dom64 / FP64_PEAK + dom32 / FP32_PEAK) if dom_fl > 0 else FP64_PEAK
        achieved = dom_fl / (means[dom] * 1e-3)
        # whole-step fractions: algorithmic (reference counts) and executed (the kernels' counters)
        n32 = np.zeros(625)
        for c in range(625):
            n32[c] = float(ck32[256 + c]) / max(1, ck[256 + c]) * NK[c] if ck[256 + c] else 0.0
        for idx in range(25):
            if cj[idx]:
                f32 = cj32[idx] / cj[idx]
                for c in range(625):
                    L1, L2, L3, L4 = cls_of(c)
                    if (L1 + L2) * 5 + (L3 + L4) == idx:
                        n32[c] += f32 * NJ[c]
        t_exec = 0.0
        for idx in range(25):
            if cj[idx]:
                L12, L34 = idx // 5, idx % 5
                cs = [c for c in range(625) if sum(cls_of(c)[:2]) == L12 and sum(cls_of(c)[2:]) == L34 and NJ[c]]
                if cs:
                    Fm = sum(NJ[c] * cls_flops(c) for c in cs) / max(1, sum(NJ[c] for c in cs))
                    t_exec += (cj[idx] - cj32[idx]) * Fm / FP64_PEAK + cj32[idx] * Fm / FP32_PEAK
        for c in range(625):
            if ck[256 + c]:
                t_exec += (ck[256 + c] - ck32[256 + c]) * cls_flops(c) / FP64_PEAK + ck32[256 + c] * cls_flops(c) / FP32_PEAK
        roofline = {"bound": "valu", "pipe": "%s vector ALU (MI355X peak; the ERI core is not a dense contraction, "
                                              "no MFMA)" % ("fp32+fp64" if dom32 > 0 else "fp64"),
                    "kernel": launch_name(*dom), "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": None, "launch_ms": means[dom], "flops_per_launch": float(dom_fl),
                    "quartets_per_launch": float(Nd),
                    "flops_source": "lsjk_count_quartets (reference convention); a K launch of a class with a distinct "
                                    "mirror class counts N_K(class) + N_K(mirror)",
                    "launch_timing": "HIP events around each launch on its stream, class launches serialised "
                                     "(%d steps after the timed region)" % args.attribution_steps,
                    "jk_frac": wl.ideal_time(n32) / t_route,
                    "jk_exec_frac": t_exec / t_route}
        # HBM bytes per launch of the dominant kernel, from the committed counter summary of this
        # workload (two rocprofv3 passes over `bench.py --pmc-pass`, calibrated on copy_bytes)
        traffic_md = os.path.join(ROOT, "profiles", "jk_%s_hbm_traffic.md" % args.workload)
        if os.path.exists(traffic_md):
            for line in open(traffic_md):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 5 and cells[0] == roofline["kernel"]:
                    roofline["traffic"] = float(cells[4]) * 1e6
                    roofline["traffic_source"] = os.path.relpath(traffic_md, ROOT)
        return roofline

    roofline = attribute(head, t_step, args.table)
    # the FP64 route's dominant kernel as well, when another route holds the headline
    roofline_fp64 = None
    if head != "fp64":
        roofline_fp64 = attribute("fp64", timing["fp64"],
                                  (os.path.splitext(args.table)[0] + "_fp64.md") if args.table else None)

    route_info = {}
    for name in routes:
        d = dict(acc[name])
        d["ms_per_step"] = 1000 * timing[name]
        d["quartets_per_s"] = quartets / timing[name]
        d["speedup_vs_fp64"] = timing["fp64"] / timing[name]
        d["meets_accuracy"] = bool(meets(name))
        route_info[name] = d
    config = {"workload": args.workload, "basis": wl.bname, "nao": bas.nao, "nprim": bas.nprim,
              "npair": pl.npair, "threpq": 1.0e-12, "thresp": thresp, "route": head,
              "thredp": routes[head][0], "N_J": int(NJ.sum()), "N_K": int(NK.sum()), "routes": route_info,
              "pairlist_s": wl.t_pl}
    if tune_info is not None:
        config["autotune"] = tune_info
    if screen is not None:
        config["screen"] = {"model": os.path.relpath(screen_path, ROOT), "certified_classes": len(screen.certified()),
                            "trained_units": int(screen.report.get("n_train", 0))}
    head_fp32 = acc[head]["fp32_share"]
    out = {"metric": METRIC, "value": quartets / t_step, "unit": "ERI quartets/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000 * t_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32" if head_fp32 > 0 else "f64",
           "data": "synthetic", "config": config, "roofline": roofline}
    if roofline_fp64 is not None:
        out["roofline_fp64_route"] = roofline_fp64
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if not args.quick:
        others = {}
        others["c20h42_ccpvdz"] = measure_extra(ls, W, "c20h42_ccpvdz", 3,
                                                0 if args.no_cpu_baseline else 16, args.thredp,
                                                not args.no_autotune)
        others["h2o256_631gs"] = measure_extra(ls, W, "h2o256_631gs", 2,
                                               0 if args.no_cpu_baseline else 4096, args.thredp,
                                               not args.no_autotune)
        others["water_batch_256"] = measure_batch(ls, W, 256, 2, args.thredp, not args.no_autotune)
        config["other_workloads"] = others
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
