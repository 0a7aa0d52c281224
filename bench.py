"""Widom insertion throughput on one MI355X.

Workload (one step): CO2 Widom insertion with LJ + Ewald into the synthetic hexagonal
framework of the benchmark's CO2 case -- a = b = 13 Å, c = 15 Å, gamma = 120 deg, 300 random
atoms (C/H/O/Zn, 1.5 Å separation, Gaussian charges sd 0.5 e shifted to neutral), replicated
3x3x3 (8,100 atoms); real-space and LJ cutoff 12 Å, Ewald precision 1e-6 (184 coupled
k-vectors), T = 298.15 K, 1e6 insertions per step, seed 42. Inputs are resident on the device
before timing starts. The mixed path is two-phase, as the reference's own `widom`: a rigorous
hard-core bound first rejects the insertions whose Boltzmann weight is exactly 0.0 in FP64
(88.7 % on this framework), and only the rest get an energy; the FP64 plain path evaluates
every insertion.

Prints ONE JSON line. `value` is the mixed-precision path's insertions/s; the FP64 plain
path is timed in the same run on the same insertion streams. `roofline` prices the energy
kernel's algorithmic FLOPs (61 per in-cutoff charged pair, 11 per (k-vector, site), 4 per
k-vector) against the FP32 vector peak, from HIP-event kernel durations. `cpu_baseline` times
the FP64 CPU oracle (a port of the reference's CPU path) on the host's cores on a bounded
sample of the same workload. `grad_cost` is the time of widom_grad over widom on the
64-framework gradient workload at equal insertions. `parity_*` give |d ln K_H| and |d q_st|
between the mixed and FP64 paths on identical insertion streams, against the gates 1e-4 and
0.01 kJ/mol.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "pureadsorb.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

import pureadsorb_amd as pa  # noqa: E402
from pureadsorb_amd import synth  # noqa: E402

METRIC = "Widom insertions/s on 1 GPU, mixed vs FP64 at |d lnK_H|<=1e-4; grad cost"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 peak
PEAK_HBM_GBS = 8000.0
T_K = 298.15
GATE_D_LN_KH = 1e-4
GATE_D_QST_EV = 0.01e3 / 6.02214076e23 / 1.602176634e-19  # 0.01 kJ/mol


def host_cpu():
    """CPU model, logical CPUs the OS reports and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def gates(mixed, f64):
    """|d ln K_H| and |d q_st| (eV) between two WidomResults and whether each meets its gate."""
    dl, dq = abs(mixed.ln_K_H - f64.ln_K_H), abs(mixed.q_st - f64.q_st)
    return {"abs_d_ln_K_H": dl, "d_ln_K_H_gate": GATE_D_LN_KH, "d_ln_K_H_pass": bool(dl <= GATE_D_LN_KH),
            "abs_d_q_st_eV": dq, "d_q_st_gate_eV": GATE_D_QST_EV, "d_q_st_pass": bool(dq <= GATE_D_QST_EV)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload():
    fw = synth.config1_framework(42)
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    return fw, ff, g, ew


def porous_workload():
    """Config 1's cell/replication/guest/Ewald at 80 atoms per cell: about 1% of insertions
    bind, so ln K_H carries information (the specified 300-atom cell admits no insertion)."""
    _, ff, g, ew = workload()
    fw = pa.replicate(synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 80, 42, charge_sd=0.5), (3, 3, 3))
    return fw, ff, g, ew


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "widom_binned_config1_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per 1e6-insertion launch of the binned kernel on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a calibration copy)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def time_steps(b, g, ninsert, steps, warmup, precision, seed=42):
    for i in range(warmup):
        pa.widom(b, g, T=T_K, ninsert=ninsert, seed=seed + i, precision=precision)
    pa.synchronize()
    energy_ms, launches = 0.0, 0
    res = None
    t0 = time.perf_counter()
    for i in range(steps):
        res = pa.widom(b, g, T=T_K, ninsert=ninsert, seed=seed, precision=precision)
        t = pa.last_timing(b)
        energy_ms += t["energy_ms"]
        launches += t["energy_launches"]
    pa.synchronize()
    wall = time.perf_counter() - t0
    return wall, energy_ms, launches, res[0]


PAIR_SAMPLE = 20000


def cpu_baseline(fw, ff, g, ew, sample, b=None):
    """The FP64 CPU oracle's two-phase path (phase-0 rejection, then the energies of the
    survivors: the reference's CPU widom, src/widom.jl:390-420) on `sample` insertions of the
    workload, all host threads (<= 16). Also returns, untimed, the mean exact in-cutoff pair count
    over the insertions the device's phase 0 lets through and the device's rejected fraction,
    both on the first PAIR_SAMPLE insertions."""
    import oracle as O

    types = [pa.typeindex(ff, s + "_") for s in fw.symbols]
    o = O.OracleSystem(fw.cell, fw.frac, types, fw.charges, fw.replication, ff.sigma, ff.epsilon, ff.cutoff, ff.tail,
                       g.sites, g.types, g.charges, ew.cutoff, ew.precision)
    # the OpenMP team: every CPU this process may use, at most 16 (a GPU box's CPU share per GPU)
    threads = max(1, min(16, host_cpu()[2]))
    kT = pa.KB * T_K
    _, rho2 = o.rejection_tables(kT, 2.220446049250313e-16)
    o.energies_twophase(42, 0, 0, 64, rho2, nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    o.energies_twophase(42, 0, 0, sample, rho2, nthreads=threads)
    wall = time.perf_counter() - t0
    # the pair census (single-phase FP64, every insertion) on the first PAIR_SAMPLE insertions
    npair = min(sample, PAIR_SAMPLE)
    _, pairs = o.energies(42, 0, 0, npair, want_pairs=True, nthreads=threads)
    if b is not None:
        keep = pa.hardcore_flags(b, 0, 42, T_K, 0, npair) == 0
    else:
        keep = np.ones(npair, dtype=bool)
    return sample / wall, threads, float(np.mean(pairs[keep])), float(1.0 - keep.mean())


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ninsert", type=int, default=1_000_000)
    ap.add_argument("--f64-steps", type=int, default=2, help="timed steps of the FP64 plain path (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000,
                    help="insertions timed on the CPU oracle, two-phase, all host threads (0: skip)")
    ap.add_argument("--no-parity", action="store_true", help="skip the porous-variant parity line")
    ap.add_argument("--grad-steps", type=int, default=2,
                    help="timed steps of the gradient-cost measurement on the 64-framework workload (0: skip)")
    ap.add_argument("--workload", choices=("config1", "config2", "config3", "config4", "rubtak", "nvt"),
                    default="config1",
                    help="config1: the 8,100-atom CO2 case (the metric); config2: 256 frameworks x 1e5 insertions; "
                         "config3: gradient cost on 64 frameworks x 2e5 insertions; config4: GCMC move path; "
                         "rubtak: the reference's own benchmark system (RUBTAK 3x3x3 + CO2, TraPPE, 12 A); "
                         "nvt: NVT move cost and Widom insertion along the chain (RUBTAK 3x3x3 + 50 CO2)")
    ap.add_argument("--bin-scale", type=float, default=0.0,
                    help="mixed-path bin edge multiplier (0: the library tunes it on the first call)")
    ap.add_argument("--gcmc-warmup", type=int, default=20000)
    ap.add_argument("--gcmc-production", type=int, default=50000)
    ap.add_argument("--nvt-moves", type=int, default=20000, help="nvt: translation attempts per chain count, in total")
    ap.add_argument("--nvt-widom", type=int, default=256, help="nvt: test insertions per chain per cycle")
    ap.add_argument("--nvt-cycles", type=int, default=20, help="nvt: production cycles per timed run_nvt")
    ap.add_argument("--nvt-cpu-sample", type=int, default=20000, help="nvt: test insertions timed on the CPU oracle")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one GPU, and --gpus {args.gpus} is not run. "
              "Widom insertions shard by (system, insertion block) with one small reduction of per-block "
              "log-sum-exp partials.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy_mib} MiB three times")
        return

    log("device:", pa.device_name())
    if args.workload == "config2":
        return bench_config2(args)
    if args.workload == "config3":
        return bench_config3(args)
    if args.workload == "config4":
        return bench_config4(args)
    if args.workload == "rubtak":
        return bench_rubtak(args)
    if args.workload == "nvt":
        return bench_nvt(args)
    fw, ff, g, ew = workload()
    t0 = time.perf_counter()
    b = pa.FrameworkBatch([fw], ff, g, ew)
    log(f"batch: {fw.natoms} atoms, {b.nk[0]} coupled k-vectors, built in {time.perf_counter() - t0:.2f} s")
    if args.bin_scale > 0:
        b.bin_scale = args.bin_scale

    wall, ems, nl, r_mx = time_steps(b, g, args.ninsert, args.steps, args.warmup, "mixed")
    ins_s = args.steps * args.ninsert / wall
    avg_launch_ms = ems / max(nl, 1)
    ins_per_launch = args.steps * args.ninsert / max(nl, 1)
    log(f"mixed: {ins_s:.4e} ins/s, {wall / args.steps * 1e3:.3f} ms/step, energy kernel {avg_launch_ms:.3f} ms/launch, "
        f"bin scale {b.bin_scale}")

    out = {
        "metric": METRIC,
        "value": ins_s,
        "unit": "insertions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "CO2 Widom LJ+Ewald, hex cell 13/13/15 A gamma=120, 300 atoms x 3x3x3 = 8100 atoms, "
                        "rc 12 A, Ewald 1e-6 (184 coupled k), T 298.15 K, 1e6 insertions/step",
            "ninsert_per_step": args.ninsert,
            "precision": "FP32 pair + k-space terms, FP64 accumulation, tables and Boltzmann sums",
            "bin_scale": b.bin_scale,
            "parallelism": "1 GPU",
        },
    }

    f64 = None
    if args.f64_steps > 0:
        w64, e64, n64, r64 = time_steps(b, g, args.ninsert, args.f64_steps, 1, "f64")
        f64 = args.f64_steps * args.ninsert / w64
        log(f"f64 plain: {f64:.4e} ins/s, energy kernel {e64 / max(n64, 1):.3f} ms/launch")
        rel = abs(r_mx.ln_K_H - r64.ln_K_H) / abs(r64.ln_K_H)
        out["f64_plain"] = {"value": f64, "unit": "insertions/s", "ms_per_step": w64 / args.f64_steps * 1e3,
                            "energy_kernel_ms_per_launch": e64 / max(n64, 1)}
        out["speedup_mixed_vs_f64_plain"] = ins_s / f64
        out["parity_workload"] = {
            "note": "no insertion binds in this framework (min dU ~ +111 eV): ln K_H is finite only through the "
                    "log-sum-exp sums and is set by the least repulsive pose, so the absolute gate is not met "
                    "here; it is met where ln K_H carries information (porous variant, RUBTAK)",
            "ln_K_H_mixed": r_mx.ln_K_H, "ln_K_H_f64": r64.ln_K_H, "rel_d_ln_K_H": rel, **gates(r_mx, r64),
        }

    if not args.no_parity:
        pfw, pff, pg, pew = porous_workload()
        pb = pa.FrameworkBatch([pfw], pff, pg, pew)
        pm = pa.widom(pb, pg, T=T_K, ninsert=args.ninsert, seed=42, precision="mixed")[0]
        p64 = pa.widom(pb, pg, T=T_K, ninsert=args.ninsert, seed=42, precision="f64")[0]
        out["parity_porous_variant"] = {
            "workload": "same cell, replication, guest and Ewald settings, 80 atoms per cell (2160 atoms)",
            "ninsert": args.ninsert,
            **gates(pm, p64),
            "K_H_mixed": pm.K_H, "K_H_f64": p64.K_H, "q_st_mixed_eV": pm.q_st, "q_st_f64_eV": p64.q_st,
        }
        log(f"porous parity: d ln K_H {abs(pm.ln_K_H - p64.ln_K_H):.2e}, d q_st {abs(pm.q_st - p64.q_st):.2e} eV")

    # phase 0 on the timed insertions (every step uses seed 42): the energy kernel sees only survivors
    rej = float(pa.hardcore_flags(b, 0, 42, T_K, 0, args.ninsert).mean())
    out["phase0_rejected_fraction"] = rej
    log(f"phase 0 rejects {rej:.4f} of the insertions")
    pin = None
    if args.cpu_sample > 0:
        cpu_rate, threads, pin, rej_sample = cpu_baseline(fw, ff, g, ew, args.cpu_sample, b)
        model, nproc, usable = host_cpu()
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "insertions/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_sample} insertions of the same workload (seed 42, system 0; the GPU step times "
                      f"1e6 seed-42 insertions), FP64 oracle (C restatement of the reference CPU path, two-phase "
                      f"with hard-core rejection), OpenMP over {threads} threads, {args.cpu_sample / cpu_rate:.1f} s",
            "cpu_model": model, "nproc": nproc, "usable_cpus": usable,
        }
        log(f"cpu oracle: {cpu_rate:.4e} ins/s on {threads} threads; mean in-cutoff pairs of survivors {pin:.1f}")
    if pin is not None:
        N, K = g.nsites, int(b.nk[0])
        # algorithmic FLOP of one surviving insertion; rejected insertions count no FLOP
        f_ins = 61.0 * pin + 11.0 * N * K + 4.0 * K
        achieved = f_ins * ins_per_launch * (1.0 - rej) / (avg_launch_ms * 1e-3) / 1e12
        traffic = measured_traffic() if args.ninsert == 1_000_000 else None
        out["roofline"] = {
            "bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
            "kernel": "widom_binned_kernel<3,true,0>",
            "flops_per_surviving_insertion": f_ins, "in_cutoff_pairs_per_surviving_insertion": pin,
            "surviving_fraction": 1.0 - rej,
            "kernel_ms_per_launch": avg_launch_ms, "insertions_per_launch": ins_per_launch,
        }
    if args.grad_steps > 0:
        tw, tg, gms, nweighted, _ = grad_cost(args.grad_steps, 1)
        out["grad_cost"] = {
            "workload": "64 frameworks (first 64 of the 256-framework draw), CO2 LJ+Ewald rc 12 A, 2e5 insertions "
                        "each; d ln K_H and d q_st wrt per-type eps, sigma and per-atom charges, mixed path",
            "widom_ms": tw * 1e3, "widom_grad_ms": tg * 1e3, "ratio": tg / tw, "grad_kernels_ms": gms,
            "weighted_insertions": nweighted,
        }
        log(f"grad cost: widom {tw * 1e3:.1f} ms, widom_grad {tg * 1e3:.1f} ms (x{tg / tw:.2f})")
    print(json.dumps(out), flush=True)


def bench_config2(args):
    """256 synthetic frameworks (200-6000 atoms per cell, log-uniform; cells 10-40 Å, angles
    70-110 deg; replicated to widths >= 24 Å), CO2 LJ + Ewald, 1e5 insertions each in one call."""
    fws = synth.config2_frameworks(256)
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    t0 = time.perf_counter()
    b = pa.FrameworkBatch(fws, ff, g, ew)
    build_s = time.perf_counter() - t0
    ninsert = 256 * 100_000
    log(f"config2 batch: {sum(f.natoms for f in fws)} atoms in 256 frameworks, built in {build_s:.1f} s")
    wall, ems, nl, _ = time_steps(b, g, ninsert, args.steps, args.warmup, "mixed")
    rm = pa.widom(b, g, T=T_K, ninsert=ninsert, seed=42, precision="mixed")
    t64 = time.perf_counter()
    r64 = pa.widom(b, g, T=T_K, ninsert=ninsert, seed=42, precision="f64")
    w64 = time.perf_counter() - t64
    ins_s = args.steps * ninsert / wall
    dln = np.array([abs(a.ln_K_H - c.ln_K_H) for a, c in zip(rm, r64)])
    rel = dln / np.maximum(1.0, np.abs([c.ln_K_H for c in r64]))
    out = {
        "metric": METRIC + " (256-framework batch)", "value": ins_s, "unit": "insertions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 frameworks, 200-6000 atoms/cell log-uniform, cells 10-40 A, angles 70-110, "
                               "widths >= 24 A, CO2 LJ+Ewald rc 12 A, 1e5 insertions each", "ninsert_per_step": ninsert},
        "batch_build_s": build_s,
        "f64_plain": {"value": ninsert / w64, "unit": "insertions/s"},
        "speedup_mixed_vs_f64_plain": ins_s / (ninsert / w64),
        "energy_kernel_ms_per_step": ems / args.steps,
        "K_H_mixed": [r.K_H for r in rm], "K_H_f64": [r.K_H for r in r64],
        "ln_K_H_mixed": [r.ln_K_H for r in rm], "ln_K_H_f64": [r.ln_K_H for r in r64],
        "max_abs_d_ln_K_H": float(dln.max()), "max_rel_d_ln_K_H": float(rel.max()),
    }
    # systems whose ln K_H shows binding insertions (above -50) against the |d ln K_H| <= 1e-4
    # gate; the others relative to |ln K_H|
    lnf = np.array([c.ln_K_H for c in r64])
    bind = lnf > -50.0
    out["binding_systems"] = int(bind.sum())
    out["binding_max_abs_d_ln_K_H"] = float(dln[bind].max()) if bind.any() else None
    out["nonbinding_max_rel_d_ln_K_H"] = float(rel[~bind].max()) if (~bind).any() else None
    log(f"config2 mixed: {ins_s:.4e} ins/s; f64 {ninsert / w64:.4e}; max |d ln K_H| {dln.max():.2e}")
    out["roofline"] = config2_roofline(fws, ff, g, ew, b, ems / args.steps)
    log(f"config2 roofline: {out['roofline']['achieved']:.2f} TFLOP/s = {100 * out['roofline']['frac']:.1f} % of "
        f"{PEAK_FP32_TFLOPS}")
    print(json.dumps(out), flush=True)


def config2_roofline(fws, ff, g, ew, b, energy_ms):
    """Roofline of the 256-framework call's energy kernel: the algorithmic FLOPs of the surviving
    insertions (61 per in-cutoff pair, 11 per (k-vector, site), 4 per k-vector) over the energy
    kernel time per call. Survivors per framework from the device's phase 0 on the timed streams
    (seed 42, all 1e5 insertions); in-cutoff pairs per survivor from the FP64 oracle on 64
    insertions of every 4th framework, extended to the others by the survivor-weighted mean
    (untimed; about 15 s of host threads)."""
    import oracle as O

    nsys, per = len(fws), 100_000
    surv = np.array([1.0 - pa.hardcore_flags(b, s, 42, T_K, 0, per).mean() for s in range(nsys)])
    K = np.array([int(b.nk[b.framework_of[s]]) for s in range(nsys)], dtype=np.float64)
    threads = max(1, min(16, host_cpu()[2]))
    w_sum, wp_sum, sample = 0.0, 0.0, list(range(0, nsys, 4))
    for s in sample:
        fw = fws[s]
        types = [pa.typeindex(ff, x + "_") for x in fw.symbols]
        o = O.OracleSystem(fw.cell, fw.frac, types, fw.charges, fw.replication, ff.sigma, ff.epsilon, ff.cutoff,
                           ff.tail, g.sites, g.types, g.charges, ew.cutoff, ew.precision)
        _, pairs = o.energies(42, s, 0, 64, want_pairs=True, nthreads=threads)
        keep = pa.hardcore_flags(b, s, 42, T_K, 0, 64) == 0
        if keep.any():
            w_sum += surv[s]
            wp_sum += surv[s] * float(np.mean(pairs[keep]))
    pairs_per_survivor = wp_sum / w_sum if w_sum > 0 else 0.0
    flops = per * float(np.sum(surv * (61.0 * pairs_per_survivor + 11.0 * g.nsites * K + 4.0 * K)))
    achieved = flops / (energy_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "widom_binned_kernel<3,true,0>",
            "flops_per_call": flops, "surviving_fraction": float(surv.mean()),
            "in_cutoff_pairs_per_surviving_insertion": pairs_per_survivor, "mean_k_vectors": float(K.mean()),
            "energy_kernel_ms_per_call": energy_ms,
            "pairs_sample": f"64 insertions of each of {len(sample)} frameworks (every 4th), FP64 oracle, {threads} threads"}


def bench_rubtak(args):
    """The reference's published Widom benchmark system (RUBTAK 3x3x3, 3,078 atoms, CO2, TraPPE,
    LJ 12 A / Ewald 12 A at precision 1e-6, 298.15 K, 1e6 insertions per call): mixed and FP64
    plain throughput on the same streams, for comparison with the reference's own numbers on
    other GPUs (docs/src/benchmarks.md)."""
    golden = os.path.join(ROOT, "tests", "golden")
    ff = pa.read_forcefield(os.path.join(golden, "trappe.yaml"))
    g = pa.read_guest(os.path.join(golden, "co2.yaml"), ff)
    fw = pa.replicate(pa.read_cif(os.path.join(golden, "RUBTAK.cif")), (3, 3, 3))
    b = pa.FrameworkBatch([fw], ff, g, pa.EwaldParams(12.0, 1e-6))
    wall, ems, nl, rm = time_steps(b, g, args.ninsert, args.steps, args.warmup, "mixed")
    w64, e64, n64, r64 = time_steps(b, g, args.ninsert, max(1, args.f64_steps), 1, "f64")
    rej = float(pa.hardcore_flags(b, 0, 42, T_K, 0, min(args.ninsert, 200_000)).mean())
    out = {"metric": "Widom insertions/s on 1 GPU (RUBTAK + CO2)", "value": args.steps * args.ninsert / wall,
           "unit": "insertions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "RUBTAK.cif (tests/golden), replicated 3x3x3",
           "config": {"workload": "RUBTAK 3x3x3 (3,078 atoms) + CO2, TraPPE, LJ 12 A, Ewald 12 A / 1e-6, 298.15 K",
                      "ninsert_per_step": args.ninsert},
           "f64_plain": {"value": max(1, args.f64_steps) * args.ninsert / w64, "unit": "insertions/s"},
           "energy_kernel_ms_per_launch": ems / max(nl, 1), "phase0_rejected_fraction": rej,
           "ln_K_H_mixed": rm.ln_K_H, "ln_K_H_f64": r64.ln_K_H, "q_st_mixed_eV": rm.q_st, "q_st_f64_eV": r64.q_st}
    log(f"rubtak: mixed {out['value']:.4e} ins/s, f64 {out['f64_plain']['value']:.4e} ins/s, "
        f"d ln K_H {abs(rm.ln_K_H - r64.ln_K_H):.2e}")
    print(json.dumps(out), flush=True)


def bench_nvt(args):
    """The reference's NVT benchmark system (RUBTAK 3x3x3, 3,078 atoms, + 50 CO2, TraPPE, LJ 12 A,
    Ewald 12 A / 1e-6, 298.15 K, FP64; docs/src/benchmarks.md:55-75): translation-move cost per
    attempt at 1, 64, 256 and 1,024 chains (the reference times one mc_step! launch per attempt,
    this library runs every chain's attempts back to back in one launch), then run_nvt with
    Widom insertion along the chain on 64 chains. `value` is the test insertions/s of the
    chain-insertion kernel (HIP events); its roofline prices FP64 FLOPs (29 per (site, atom)
    distance test, 88 per in-cutoff charged (site, atom) pair, 10 per (k-vector, site) and 5 per
    k-vector, 29 per guest-guest site-pair test) against the FP64 peak."""
    import oracle as O

    golden = os.path.join(ROOT, "tests", "golden")
    ff = pa.read_forcefield(os.path.join(golden, "trappe.yaml"))
    g = pa.read_guest(os.path.join(golden, "co2.yaml"), ff)
    fw = pa.replicate(pa.read_cif(os.path.join(golden, "RUBTAK.cif")), (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    ref_us = {1: 187.2, 64: 18.2, 256: 15.8, 1024: 10.8}  # RTX 4070, FP64, docs/src/benchmarks.md:64-68
    moves = []
    for nsys in (1, 64, 256, 1024):
        b = pa.FrameworkBatch([fw] * nsys, ff, g, ew, fullk=True)
        t0 = time.perf_counter()
        st = pa.SystemState(b, g, [50] * nsys, T=T_K, seed=1)
        t_state = time.perf_counter() - t0
        nm = max(200, args.nvt_moves // nsys)
        pa.mc_moves(st, pa.MOVE_TRANSLATION, 10, T=T_K, step_trans=0.3)
        pa.synchronize()
        t0 = time.perf_counter()
        pa.mc_moves(st, pa.MOVE_TRANSLATION, nm, T=T_K, step_trans=0.3)
        pa.synchronize()
        wall = time.perf_counter() - t0
        acc = sum(st.info(n)["accepted"][0] for n in range(nsys)) / (nsys * (nm + 10))
        moves.append({"chains": nsys, "attempts_per_chain": nm, "us_per_move": wall / (nm * nsys) * 1e6,
                      "wall_s": wall, "translation_acceptance": acc, "state_build_s": t_state,
                      "reference_rtx4070_f64_us_per_move": ref_us[nsys]})
        log(f"nvt moves: {nsys} chains, {nm} attempts each: {wall / (nm * nsys) * 1e6:.3f} us/move "
            f"(reference RTX 4070 FP64 {ref_us[nsys]} us), acceptance {acc:.3f}")
        st.close()
        b.close()
    # run_nvt with Widom insertion along the chain
    nsys, nw = 64, args.nvt_widom
    b = pa.FrameworkBatch([fw] * nsys, ff, g, ew, fullk=True)
    st = pa.SystemState(b, g, [50] * nsys, T=T_K, seed=3)
    kw = dict(T=T_K, n_warmup=2, n_widom_per_cycle=nw, n_audit=1_000_000, step_trans=0.3, step_rot=0.3,
              widom_seed=1003, nblocks=5)
    pa.run_nvt(b, st, g, ff, n_production=10, **kw)
    ins_ms, ins_n, mv_ms, tot_ms, ncyc = 0.0, 0, 0.0, 0.0, 0
    res = None
    for _ in range(args.steps):
        res = pa.run_nvt(b, st, g, ff, n_production=args.nvt_cycles, **kw)
        t = pa.nvt_last_timing(st)
        ins_ms += t["insertion_ms"]
        ins_n += t["insertions"]
        mv_ms += t["move_ms"]
        tot_ms += t["total_ms"]
        ncyc += args.nvt_cycles
    ins_s = ins_n / (ins_ms * 1e-3)
    # algorithmic FP64 FLOPs per test insertion; in-cutoff (site, atom) pairs from the oracle on a
    # sample of the same insertion stream (system 0)
    types = [pa.typeindex(ff, s_ + "_") for s_ in fw.symbols]
    osys = O.OracleSystem(fw.cell, fw.frac, types, fw.charges, fw.replication, ff.sigma, ff.epsilon, ff.cutoff,
                          ff.tail, g.sites, g.types, g.charges, ew.cutoff, ew.precision)
    _, npairs = osys.energies(1003, 0, 0, 2000, want_pairs=True, nthreads=16)
    p_in = float(np.mean(npairs))
    ns_, na, nk = len(g.charges), fw.natoms, int(b.nk[0])
    flops = 29.0 * ns_ * na + 88.0 * p_in + (10.0 * ns_ + 5.0) * nk + 29.0 * 50 * ns_ * ns_
    achieved = flops * ins_s / 1e12
    # CPU baseline: the oracle's chain insertion (FP64 port) on the host cores, same workload
    chain = O.OracleMC(osys, 50, 1, *st.guests(0)[:2])
    ncpu = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    chain.chain_energies(1003, 0, 0, args.nvt_cpu_sample, nthreads=ncpu)
    cpu_s = time.perf_counter() - t0
    out = {"metric": "NVT test insertions/s into loaded configurations on 1 GPU (FP64)", "value": ins_s,
           "unit": "insertions/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": tot_ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "RUBTAK.cif (tests/golden), replicated 3x3x3; 50 CO2 per chain",
           "config": {"workload": f"run_nvt on {nsys} chains of RUBTAK 3x3x3 + 50 CO2, TraPPE, LJ 12 A, Ewald 12 A "
                                  f"/ 1e-6, 298.15 K, {args.nvt_cycles} production cycles of 50 attempts, {nw} test "
                                  "insertions per chain per cycle", "chains": nsys, "n_widom_per_cycle": nw},
           "roofline": {"bound": "fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                        "flops_per_insertion": flops, "in_cutoff_pairs_per_insertion": p_in, "k_vectors": nk},
           "cpu_baseline": {"value": args.nvt_cpu_sample / cpu_s, "unit": "insertions/s", "cores": ncpu,
                            "kind": "port", "sample": f"{args.nvt_cpu_sample} test insertions into chain 0's initial "
                                                     f"50-guest configuration, FP64 oracle, {cpu_s:.1f} s"},
           "move_cost": moves,
           "nvt_run": {"cycles": ncyc, "move_kernel_ms": mv_ms, "insertion_kernel_ms": ins_ms, "wall_ms": tot_ms,
                       "moves_per_s": ncyc * 50 * nsys / (mv_ms * 1e-3),
                       "acceptance": list(res[0].acceptance), "mu_ex_eV": res[0].mu_ex, "q_st_eV": res[0].q_st,
                       "energy_eV": res[0].energy}}
    log(f"nvt: {ins_s:.4e} test insertions/s ({achieved:.2f} TFLOP/s FP64, {achieved / PEAK_FP64_TFLOPS:.1%}), "
        f"moves {ncyc * 50 * nsys / (mv_ms * 1e-3):.3e}/s, cpu {args.nvt_cpu_sample / cpu_s:.3e}/s")
    print(json.dumps(out), flush=True)


def grad_cost(steps, warmup):
    """Gradient cost: the first 64 frameworks of the 256-framework draw, CO2, 2e5 insertions each;
    wall time of widom_grad (weights, gradient kernels, host assembly) over widom's at equal
    insertions. Returns (widom s, widom_grad s, gradient-kernel ms, weighted insertions, results)."""
    fws = synth.config2_frameworks(256)[:64]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    b = pa.FrameworkBatch(fws, ff, g, ew)
    ninsert = 64 * 200_000
    log(f"config3 batch: {sum(f.natoms for f in fws)} atoms in 64 frameworks")
    for _ in range(warmup):
        pa.widom(b, g, T=T_K, ninsert=ninsert, seed=42, precision="mixed")
        pa.widom_grad(b, g, T=T_K, ninsert=ninsert, seed=42, precision="mixed")
    pa.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pa.widom(b, g, T=T_K, ninsert=ninsert, seed=42, precision="mixed")
    pa.synchronize()
    tw = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    gms = 0.0
    for _ in range(steps):
        gr = pa.widom_grad(b, g, T=T_K, ninsert=ninsert, seed=42, precision="mixed")
        gms += pa.last_timing(b)["grad_ms"]
    pa.synchronize()
    tg = (time.perf_counter() - t0) / steps
    return tw, tg, gms / steps, pa.last_timing(b)["grad_insertions"], gr


def bench_config3(args):
    """Gradient cost on the 64-framework workload (grad_cost) as its own JSON line."""
    ninsert = 64 * 200_000
    tw, tg, gms, nweighted, gr = grad_cost(args.steps, args.warmup)
    gms *= args.steps
    out = {
        "metric": METRIC + " (gradient cost, 64-framework batch)", "value": ninsert / tg,
        "unit": "insertions/s with gradients", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tg * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 frameworks (first 64 of the 256-framework draw), CO2 LJ+Ewald rc 12 A, 2e5 "
                               "insertions each; d ln K_H and d q_st wrt per-type eps, sigma and per-atom charges",
                   "ninsert_per_step": ninsert},
        "widom_ms": tw * 1e3, "widom_grad_ms": tg * 1e3, "grad_cost": tg / tw,
        "grad_kernel_ms": gms / args.steps, "weighted_insertions": nweighted,
        "charge_entries": int(sum(len(x.dlnkh_dq) for x in gr)),
    }
    log(f"config3: widom {tw * 1e3:.1f} ms, widom_grad {tg * 1e3:.1f} ms (cost x{tg / tw:.2f}), grad kernel "
        f"{gms / args.steps:.1f} ms over {nweighted} weighted insertions")
    print(json.dumps(out), flush=True)


def bench_config4(args):
    """GCMC move path: CO2 in the config-1 cell replicated 3x3x2 (the smallest supercell whose
    perpendicular widths reach 2 x 12 A; 2x2x2 gives 22.5 A), three chains at fugacity 1e4, 1e5,
    1e6 Pa, capacity 400, moves insertion/deletion/translation (Gaussian, s.d. 1 A)/rotation at
    1/4 each, 2e4 equilibration + 5e4 production cycles. Run on the specified 300-atom cell (which
    accepts no insertion) and on a 40-atom porous variant of the same cell; the first 200
    production-like cycles of each chain are replayed on the FP64 oracle from the same state and
    streams, and the run fails unless every porous-variant chain accepted insertions in them."""
    import oracle as O

    fug = [1.0e4, 1.0e5, 1.0e6]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    out = {"metric": "GCMC move attempts/s on 1 GPU (move-energy path, FP64)", "unit": "attempts/s", "n_gpus": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "steps": 1, "warmup": 0,
           "config": {"workload": "CO2 GCMC, config-1 cell (13/13/15 A, gamma 120) x 3x3x2, 298.15 K, fugacity 1e4/1e5/1e6 "
                                  "Pa, capacity 400, p(insert)=p(delete)=p(translate, s.d. 1 A)=p(rotate, step 0.5)=1/4, "
                                  f"{args.gcmc_warmup} + {args.gcmc_production} cycles, audit every 5000 cycles"}}
    variants = {"config1_cell": synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 300, 42, charge_sd=0.5),
                "porous_40_atoms": synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 40, 42, charge_sd=0.5)}
    total_attempts, total_s = 0, 0.0
    for name, uc in variants.items():
        fw = pa.replicate(uc, (3, 3, 2))
        b = pa.FrameworkBatch([fw] * 3, ff, g, ew, fullk=True)
        # oracle replay of the first cycles from an empty start
        st0 = pa.SystemState(b, g, [0, 0, 0], T=T_K, seed=7, capacities=[400] * 3)
        nrep = 200
        pa.run_gcmc(st0, T=T_K, n_warmup=0, n_production=nrep, n_audit=nrep, step_trans=1.0, step_rot=0.5,
                    fugacity=fug, nblocks=2)
        replay = []
        types = [pa.typeindex(ff, s + "_") for s in fw.symbols]
        osys = O.OracleSystem(fw.cell, fw.frac, types, fw.charges, fw.replication, ff.sigma, ff.epsilon, ff.cutoff,
                              ff.tail, g.sites, g.types, g.charges, ew.cutoff, ew.precision)
        for n in range(3):
            e_dev, n_dev = pa.gcmc_samples(st0, n, nrep)
            m = O.OracleMC(osys, 400, O.splitmix64(7, n + 1))
            e_or, n_or = [], []
            for _ in range(nrep):
                m.cycle((0.25, 0.25, 0.0, 0.5), 1, 1.0, 0.5, fug[n], T_K)
                sc = m.scalars()
                e_or.append(sc["energy"])
                n_or.append(sc["occupancy"])
            acc_o, _ = m.counts()
            replay.append({"fugacity_pa": fug[n], "cycles": nrep, "occupancy_sequences_equal": bool(np.array_equal(n_dev, n_or)),
                           "max_abs_d_energy_eV": float(np.max(np.abs(e_dev - np.array(e_or)))),
                           "oracle_accepted": [int(x) for x in acc_o], "device_accepted": st0.info(n)["accepted"]})
            if not np.array_equal(n_dev, n_or):
                raise SystemExit(f"config4 {name} chain {n}: device and oracle occupancy sequences differ")
            if name.startswith("porous") and acc_o[3] == 0:
                raise SystemExit(f"config4 {name} chain {n}: no insertion accepted in {nrep} replayed cycles")
        st0.close()
        st = pa.SystemState(b, g, [0, 0, 0], T=T_K, seed=42, capacities=[400] * 3)
        t0 = time.perf_counter()
        res = pa.run_gcmc(st, T=T_K, n_warmup=args.gcmc_warmup, n_production=args.gcmc_production, n_audit=5000,
                          step_trans=1.0, step_rot=0.5, fugacity=fug, nblocks=10)
        wall = time.perf_counter() - t0
        att = sum(sum(r.attempted) for r in res)
        # warmup attempts are not in the production counters; count them from the chain counters
        att_all = sum(st.info(n)["rng_counter"] for n in range(3))
        total_attempts += att_all
        total_s += wall
        st.close()
        out[name] = {
            "atoms": fw.natoms, "wall_s": wall, "attempts": att_all, "attempts_per_s": att_all / wall,
            "results": [{"fugacity_pa": fug[n], "loading": r.loading, "loading_err": r.loading_err,
                         "energy_eV": r.energy, "energy_err_eV": r.energy_err, "q_st_eV": r.q_st,
                         "q_st_err_eV": r.q_st_err, "max_occupancy": r.max_occupancy,
                         "accepted": r.accepted, "attempted": r.attempted} for n, r in enumerate(res)],
            "oracle_replay": replay,
        }
        log(f"config4 {name}: {att_all} attempts in {wall:.1f} s ({att_all / wall:.3e}/s); loadings "
            + ", ".join(f"{r.loading:.2f}+-{r.loading_err:.2f}" for r in res)
            + f"; replay equal: {[x['occupancy_sequences_equal'] for x in replay]}")
    out["value"] = total_attempts / total_s
    out["ms_per_step"] = total_s * 1e3
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
