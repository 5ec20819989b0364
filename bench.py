#!/usr/bin/env python3
"""Throughput of the channel time-stepper on one MI355X.

A step is one RK3 time step (three RK_CHAN_1 substeps) of the full device-
resident state.  The default workload is the 384 x 257 x 384 channel at
Re = 395 (tanh grid, CS = 1.75, dt = 5e-4, seeded perturbation 0.05 plus the
streak seed), FP64.  Inputs are resident in HBM before the timed region.

Prints one JSON line:
  value        Mpts * RK-substeps / s (NX*NY*NZ*3*steps / seconds / 1e6)
  roofline     dominant kernel: algorithmic bytes per launch / mean HIP-event
               duration of that kernel over the timed region, against 8 TB/s
  cpu_baseline the oracle (C restatement of the reference path, OpenMP) on a
               bounded sample of the same workload on the host cores
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0   # dense BF16 matrix-core peak of MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(cfg, jlo: int, s: int = 8) -> dict:
    """Algorithmic HBM bytes per launch of each kernel (inputs read once,
    outputs written once), s bytes per real."""
    nin = 3 + cfg.n_th
    p2 = cfg.ny + 2
    mk, mz = cfg.nkx + 1, cfg.tnkz + 1
    ps = mk * mz
    zrow = cfg.nz * mk
    pl = cfg.ny - jlo + 1
    c = 2 * s
    return {
        "zinv": c * (nin * p2 * ps + nin * p2 * zrow),
        "xfused": c * (nin * (pl + 2) * zrow + 2 * nin * pl * zrow),
        "zfwd_nl": c * (2 * nin * pl * zrow + nin * pl * ps),
        # explicit RHS + CN: reads CU1-3, CTH, CP, CF (4), NL (4); writes CF (4), CU1-3, CTH
        "cn": c * ((5 + 4 + 4) + (4 + 4)) * p2 * ps,
        # REM_DIV + CP update: reads CU1-3, CP and the two Poisson factors (reals); writes CU1-3, CP
        "remdiv": c * (4 + 1 + 4) * cfg.ny * ps,
        # the fused column kernel (timed as cn): reads CU1-3, CTH, CP, CF (4), NL (4) and the two
        # Poisson factors; writes CF (4), CU1-3, CTH, CP
        "colfused": c * ((5 + 4 + 4) + (4 + 5)) * p2 * ps + c * cfg.ny * ps,
    }


def substep_model_bytes(cfg, s: int = 8) -> float:
    """Stage-level byte model per substep: s*[2(3+N)P + (23+7N)F]."""
    P = cfg.nx * cfg.ny * cfg.nz
    F = 2 * (cfg.nkx + 1) * (2 * cfg.nkz + 1) * cfg.ny
    return s * (2 * (3 + cfg.n_th) * P + (23 + 7 * cfg.n_th) * F)


def time_path(cfg, phys, precision: str, steps: int, warmup: int, kernel_timing: bool, zdft: bool = False,
              tune: bool = False, graph_pass: bool = True):
    """Initial state -> warm-up -> [autotune] -> timed steps; returns (seconds,
    kernel times, channel); the autotuner's choice is in channel.tuned.

    With kernel timing the timed steps run eagerly (an event pair around each
    hot kernel on the context's stream, so no captured graphs); the same
    number of steps is then timed again with the per-stage graphs and no
    events, the production path, into channel.graph_elapsed (graph_pass=False
    skips it: a profiler's trace then holds only the event-bracketed
    launches)."""
    from diablo_amd import gpu
    g = gpu.GpuChannel(cfg, precision=precision)
    g.initialize(*phys)
    if zdft:
        g.set_zdft(True)
    g.step(warmup)
    g.synchronize()
    g.tuned = g.autotune(3) if tune else None
    if kernel_timing:
        g.set_kernel_timing(True)
        g.kernel_times()  # reset
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    kt = g.kernel_times() if kernel_timing else {}
    g.set_kernel_timing(False)
    g.graph_elapsed = None
    if kernel_timing and graph_pass:
        g.step(1)  # captures the stage graphs outside the timed steps
        g.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        g.graph_elapsed = time.perf_counter() - t0
    return elapsed, kt, g


def accuracy(cfg, phys, steps: int, zdft: bool = False) -> dict:
    """Mixed path against the FP64 path after `steps` steps from the same state."""
    from diablo_amd import gpu, metrics
    g64 = gpu.GpuChannel(cfg, precision="fp64")
    gmx = gpu.GpuChannel(cfg, precision="mixed")
    g64.initialize(*phys)
    gmx.initialize(*phys)
    if zdft:
        gmx.set_zdft(True)
    g64.step(steps)
    gmx.step(steps)
    g64.synchronize()
    gmx.synchronize()
    r = metrics.compare(g64, g64.state(), gmx, gmx.state())
    g64.close()
    gmx.close()
    r["steps"] = steps
    return r


def ensemble_run(n: int, steps: int, warmup: int, precision: str = "mixed", acc_steps: int = 0,
                 tune: bool = False) -> dict:
    """Aggregate throughput of n independent KH runs on uneven grids advanced
    together (members with the same NX, NZ share one launch per kernel kind,
    one graph per RK stage); with acc_steps > 0 also the worst member's error
    against an FP64 ensemble run from the same states."""
    from diablo_amd import case, gpu, metrics
    cfgs = case.ensemble_cases(n)
    ens = gpu.GpuEnsemble(cfgs, precision=precision)
    phys = [case.initial_physical(cfg) for cfg in cfgs]
    for p, m in zip(phys, ens.members):
        m.initialize(*p)
    ens.step(warmup)
    ens.synchronize()
    tuned = ens.autotune(2) if tune else None
    t0 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    el = time.perf_counter() - t0
    pts = ens.points
    per = [cfg.points * 3 * steps / el / 1e6 for cfg in cfgs]
    out = {"value": round(pts * 3 * steps / el / 1e6, 2), "unit": "Mpts*substeps/s", "members": n,
           "steps": steps, "precision": precision, "ms_per_step": round(el / steps * 1e3, 3),
           "points_total": pts, "per_member_min": round(min(per), 2), "per_member_max": round(max(per), 2),
           "grids": "NX=NZ in {64,96,128,192} x NY in {33,65,97,129}, Ri in [0.05,0.25], Re in {500,1000}",
           "autotune": tuned}
    ens.close()
    if acc_steps > 0 and precision != "fp64":
        e64 = gpu.GpuEnsemble(cfgs, precision="fp64")
        emx = gpu.GpuEnsemble(cfgs, precision=precision)
        for p, a, b in zip(phys, e64.members, emx.members):
            a.initialize(*p)
            b.initialize(*p)
        e64.step(acc_steps)
        emx.step(acc_steps)
        e64.synchronize()
        emx.synchronize()
        worst = {"rel_l2": 0.0, "ke_drift": 0.0, "max_div": 0.0}
        for a, b in zip(e64.members, emx.members):
            r = metrics.compare(a, a.state(), b, b.state())
            for k in worst:
                worst[k] = max(worst[k], r[k])
        e64.close()
        emx.close()
        out.update({"rel_l2_vs_fp64_worst_member": worst["rel_l2"], "ke_drift_vs_fp64_worst_member": worst["ke_drift"],
                    "max_div_worst_member": worst["max_div"], "accuracy_steps": acc_steps})
    return out


def zdft_flops(cfg) -> dict:
    """Useful (unpadded) flops per launch of the DFT-as-GEMM z kernels: real
    GEMMs of (2 NZ x 2 MZ) and (2 MZ x 4 NZ) per column block of NKX+1 modes,
    three split-BF16 products each."""
    nz, mz, mk, nin = cfg.nz, 2 * cfg.nkz + 1, cfg.nkx + 1, 3 + cfg.n_th
    jlo = min([2] + [1 if cfg.th_bc_ymin[n] == 1 else 2 for n in range(cfg.n_th)])
    inv = 3 * 2 * (2 * nz) * (2 * mz) * mk * nin * (cfg.ny + 2)
    fwd = 3 * 2 * (2 * mz) * (4 * nz) * mk * nin * (cfg.ny - jlo + 1)
    return {"zinv": float(inv), "zfwd": float(fwd)}


def zdft_run(steps: int, warmup: int, acc_steps: int) -> dict:
    """The Kelvin-Helmholtz layer at 256 x 193 x 256: FP64, mixed with FFT z
    transforms, mixed with the split-BF16 DFT-as-GEMM z transforms."""
    from diablo_amd import case
    cfg = case.WORKLOADS["kh256"]()
    phys = case.initial_physical(cfg)
    pts = cfg.points * 3 * steps
    el64, _, g = time_path(cfg, phys, "fp64", steps, warmup, False)
    g.close()
    elm, _, g = time_path(cfg, phys, "mixed", steps, warmup, False)
    g.close()
    elz, ktz, g = time_path(cfg, phys, "mixed", steps, warmup, True, zdft=True)
    g.close()
    el_nt, _, g = time_path(cfg, phys, "mixed", steps, warmup, False, zdft=True)
    g.close()
    el_t, _, g = time_path(cfg, phys, "mixed", steps, warmup, False, tune=True)
    tuned = g.tuned
    g.close()
    _, ktf, g = time_path(cfg, phys, "mixed", steps, warmup, True)
    g.close()
    fl = zdft_flops(cfg)
    # both variants move the same HBM bytes (FP32 state and intermediates);
    # a z kernel cannot beat max(bytes / HBM peak, flops / MFMA peak)
    jlo = min([2] + [1 if cfg.th_bc_ymin[n] == 1 else 2 for n in range(cfg.n_th)])
    kb = kernel_bytes(cfg, jlo, s=4)
    kern = {}
    for kind, name in (("zinv", "zinv"), ("zfwd_nl", "zfwd")):
        ms, n = ktz.get(kind, (0.0, 0))
        msf, nf = ktf.get(kind, (0.0, 0))
        if n:
            avg = ms / n
            hbm_ms = kb[kind] / (HBM_PEAK_GBS * 1e9) * 1e3
            mfma_ms = fl[name] / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e3
            kern[name] = {"ms_per_launch": round(avg, 4), "gflop_per_launch": round(fl[name] / 1e9, 2),
                          "achieved_tflops": round(fl[name] / (avg * 1e-3) / 1e12, 1),
                          "frac_of_bf16_dense_peak": round(fl[name] / (avg * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                          "hbm_bytes_per_launch": kb[kind], "hbm_floor_ms": round(hbm_ms, 4),
                          "mfma_floor_ms": round(mfma_ms, 4),
                          "fft_ms_per_launch": round(msf / nf, 4) if nf else None}
    out = {"workload": f"kh256: Kelvin-Helmholtz {cfg.nx}x{cfg.ny}x{cfg.nz}, Re=1000, Ri=0.1, dt={cfg.delta_t}",
           "fp64": round(pts / el64 / 1e6, 2), "mixed_fft": round(pts / elm / 1e6, 2),
           "mixed_zdft": round(pts / el_nt / 1e6, 2), "mixed_autotuned": round(pts / el_t / 1e6, 2),
           "autotune": tuned, "unit": "Mpts*substeps/s", "steps": steps,
           "kernels": kern}
    if acc_steps > 0:
        acc = accuracy(cfg, phys, acc_steps, zdft=True)
        out.update({"rel_l2_vs_fp64": acc["rel_l2"], "ke_drift_vs_fp64": acc["ke_drift"],
                    "max_div": acc["max_div"], "accuracy_steps": acc["steps"]})
    return out


def les_run(cfg, phys, steps: int, warmup: int) -> dict:
    """RK_CHAN_1 with LES (the constant Smagorinsky model called in every
    substep, Crank-Nicolson solves in physical space), FP64, same workload."""
    from diablo_amd import gpu
    g = gpu.GpuChannel(cfg, precision="fp64")
    g.initialize(*phys)
    g.step(1)
    g.set_les(True, True)
    g.step(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    el = time.perf_counter() - t0
    cu1 = g.get_field(gpu.CU1)
    if not math.isfinite(float(abs(cu1).max())):
        raise RuntimeError("LES state diverged")
    g.close()
    return {"value": round(cfg.points * 3 * steps / el / 1e6, 2), "unit": "Mpts*substeps/s",
            "ms_per_step": round(el / steps * 1e3, 3), "steps": steps, "dtype": "f64",
            "model": "constant Smagorinsky, les_chan every substep"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, phys, threads: int, steps: int = 3, single_core_substeps: int = 1) -> dict:
    """The oracle (C restatement of the reference path, physical-space CN as
    in the reference) on the host: `steps` RK3 steps with `threads` OpenMP
    threads from the same initial state as the GPU legs, then
    `single_core_substeps` substeps on one core (the reference's one-rank
    configuration) from the state reached."""
    from oracle import oracle
    o = oracle.Oracle(cfg, threads=threads)
    o.initialize(*phys)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    dt = time.perf_counter() - t0
    out = {"value": round(cfg.points * 3 * steps / dt / 1e6, 3), "unit": "Mpts*substeps/s", "cores": threads,
           "kind": "port",
           "sample": f"{steps} RK3 steps ({3 * steps} substeps) of the same {cfg.nx}x{cfg.ny}x{cfg.nz} case "
                     f"from the same initial state: the oracle (C restatement of RK_CHAN_1, physical-space CN "
                     f"as in the reference) with {threads} OpenMP threads, {dt:.1f} s",
           "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    if single_core_substeps > 0:
        oracle.lib().dor_set_threads(1)
        t0 = time.perf_counter()
        for k in range(single_core_substeps):
            o.rk_chan_1(1 + k % 3)
        d1 = time.perf_counter() - t0
        oracle.lib().dor_set_threads(threads)
        out["single_core"] = {"value": round(cfg.points * single_core_substeps / d1 / 1e6, 3),
                              "unit": "Mpts*substeps/s", "cores": 1,
                              "sample": f"{single_core_substeps} substep(s) on one thread, {d1:.1f} s"}
    return out


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of known size: what FETCH_SIZE / WRITE_SIZE
    read for it scales the same counters read for the solver kernels."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "repeats": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="channel384")
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--cpu-steps", type=int, default=3, help="RK3 steps of the all-core CPU baseline")
    ap.add_argument("--cpu-single-substeps", type=int, default=1,
                    help="substeps of the one-core CPU baseline (0: skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--graph-pass", choices=("on", "off"), default="on",
                    help="time the steps again with the per-stage graphs (off: only the event-timed steps run)")
    ap.add_argument("--mixed", choices=("on", "off"), default="on")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the mixed-vs-FP64 accuracy run (the 1000-step criterion)")
    ap.add_argument("--autotune", choices=("on", "off"), default="on",
                    help="pick the launch shapes with dgpu_autotune before timing (FP64 and mixed legs)")
    ap.add_argument("--les-steps", type=int, default=5, help="timed steps of the LES path (0: skip)")
    ap.add_argument("--zdft-steps", type=int, default=10,
                    help="timed steps of the 256x193x256 KH case with the MFMA z transforms (0: skip)")
    ap.add_argument("--zdft-accuracy-steps", type=int, default=2000)
    ap.add_argument("--ensemble", type=int, default=64, help="members of the batched run (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--ensemble-precision", default="mixed", choices=("mixed", "mixed_f64state", "fp64"))
    ap.add_argument("--ensemble-accuracy-steps", type=int, default=500)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB five times (reference point for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs the channel case on a single GPU "
              "(no sharding, no collectives); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)

    from diablo_amd import case, gpu

    cfg = case.WORKLOADS[args.workload]()
    log(f"[bench] workload {args.workload}: {cfg.nx}x{cfg.ny}x{cfg.nz}, building initial state")
    phys = case.initial_physical(cfg)
    log(f"[bench] FP64: warmup {args.warmup}, timing {args.steps} steps")
    elapsed, kt, g = time_path(cfg, phys, "fp64", args.steps, args.warmup, not args.no_kernel_timing,
                               tune=args.autotune == "on", graph_pass=args.graph_pass == "on")
    tuned = {"fp64": g.tuned}
    substeps = 3 * args.steps
    value = cfg.points * substeps / elapsed / 1e6
    graphs = None
    if g.graph_elapsed:
        graphs = {"note": "the same number of steps again, per-stage graphs, no kernel events",
                  "fp64": {"value": round(cfg.points * substeps / g.graph_elapsed / 1e6, 2),
                           "ms_per_step": round(g.graph_elapsed / args.steps * 1e3, 3)}}
        log(f"[bench] FP64 with graphs {graphs['fp64']['value']:.1f} Mpts*substeps/s, "
            f"{graphs['fp64']['ms_per_step']:.2f} ms/step")
    # sanity: the state must stay finite
    cu1 = g.get_field(gpu.CU1)
    if not math.isfinite(float(abs(cu1).max())):
        raise RuntimeError("state diverged")
    g.close()

    jlo = min([2] + [1 if cfg.th_bc_ymin[n] == 1 else 2 for n in range(cfg.n_th)])
    kb = kernel_bytes(cfg, jlo)
    roof = None
    if kt and kt.get("remdiv", (0, 0))[1] == 0 and kt.get("cn", (0, 0))[1] > 0:
        kb["cn"] = kb["colfused"]  # the column stage ran as the fused kernel
    if kt:
        per = {k: (ms / n if n else float("nan"), n) for k, (ms, n) in kt.items()}
        for k, (avg, n) in per.items():
            if not n:
                continue
            log(f"[bench] kernel {k:8s} launches {n:5d} mean {avg:8.3f} ms  "
                f"{kb[k] / (avg * 1e-3) / 1e9:8.1f} GB/s algorithmic")
        dom = max(kt, key=lambda k: kt[k][0])
        avg_s = per[dom][0] * 1e-3
        achieved = kb[dom] / avg_s / 1e9
        traffic = None
        tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
        if os.path.exists(tfile):
            traffic = json.load(open(tfile)).get(dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": kb[dom], "ms_per_launch": round(per[dom][0], 4)}
    model = substep_model_bytes(cfg)
    log(f"[bench] {value:.1f} Mpts*substeps/s, {elapsed / args.steps * 1e3:.2f} ms/step; stage model "
        f"{model * substeps / elapsed / 1e9:.0f} GB/s ({model * substeps / elapsed / 1e12 / 8 * 100:.1f}% of 8 TB/s)")

    mixed = None
    if args.mixed == "on":
        log(f"[bench] mixed: warmup {args.warmup}, timing {args.steps} steps")
        el_m, kt_m, gm = time_path(cfg, phys, "mixed", args.steps, args.warmup, not args.no_kernel_timing,
                                   tune=args.autotune == "on", graph_pass=args.graph_pass == "on")
        tuned["mixed"] = gm.tuned
        if graphs and gm.graph_elapsed:
            vg = cfg.points * substeps / gm.graph_elapsed / 1e6
            graphs["mixed"] = {"value": round(vg, 2), "ms_per_step": round(gm.graph_elapsed / args.steps * 1e3, 3),
                               "speedup_vs_fp64": round(vg / graphs["fp64"]["value"], 3)}
            log(f"[bench] mixed with graphs {vg:.1f} Mpts*substeps/s "
                f"({graphs['mixed']['speedup_vs_fp64']:.2f}x FP64 with graphs)")
        gm.close()
        vm = cfg.points * substeps / el_m / 1e6
        for k, (ms, n) in kt_m.items():
            if n:
                log(f"[bench] mixed kernel {k:8s} mean {ms / n:8.3f} ms")
        log(f"[bench] mixed {vm:.1f} Mpts*substeps/s ({vm / value:.2f}x FP64); accuracy over "
            f"{args.accuracy_steps} steps")
        acc = accuracy(cfg, phys, args.accuracy_steps)
        log(f"[bench] mixed vs FP64: {acc}")
        mixed = {"value": round(vm, 2), "unit": "Mpts*substeps/s", "ms_per_step": round(el_m / args.steps * 1e3, 3),
                 "speedup_vs_fp64": round(vm / value, 3), "dtype": "f32 state, transforms, products and solves; f64 low modes (I <= 2, |kz| <= 2)",
                 "rel_l2_vs_fp64": acc["rel_l2"], "ke_drift_vs_fp64": acc["ke_drift"],
                 "max_div": acc["max_div"], "accuracy_steps": acc["steps"]}

    batched = None
    if args.ensemble > 0:
        log(f"[bench] batched: {args.ensemble} KH members, {args.ensemble_precision}, {args.ensemble_steps} steps")
        batched = ensemble_run(args.ensemble, args.ensemble_steps, 1, args.ensemble_precision,
                               args.ensemble_accuracy_steps, tune=args.autotune == "on")
        log(f"[bench] batched aggregate {batched['value']:.1f} Mpts*substeps/s "
            f"({batched['value'] / value:.2f}x FP64 384); {batched}")

    les = None
    if args.les_steps > 0:
        log(f"[bench] LES: timing {args.les_steps} steps")
        les = les_run(cfg, phys, args.les_steps, 1)
        log(f"[bench] LES {les['value']:.1f} Mpts*substeps/s ({les['ms_per_step']:.1f} ms/step, "
            f"{les['value'] / value:.2f}x the DNS path)")

    zd = None
    if args.zdft_steps > 0:
        log(f"[bench] kh256 with the MFMA z transforms: {args.zdft_steps} steps, accuracy over "
            f"{args.zdft_accuracy_steps}")
        zd = zdft_run(args.zdft_steps, 1, args.zdft_accuracy_steps)
        log(f"[bench] kh256: {zd}")

    cpu = None
    if args.cpu_baseline == "on":
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: oracle, {args.cpu_steps} steps on {threads} threads, "
            f"{args.cpu_single_substeps} substep(s) on 1")
        cpu = cpu_baseline(cfg, phys, threads, args.cpu_steps, args.cpu_single_substeps)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": "Mpts·RK-substeps/s FP64 vs mixed+batched at rel-L2≤1e-5, 1 MI355X",
        "value": round(value, 2),
        "unit": "Mpts*substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar profile + perturbations, tanh y-grid)",
        "config": {"workload": f"{args.workload}: channel {cfg.nx}x{cfg.ny}x{cfg.nz}, Re={1 / cfg.nu:.0f}, "
                               f"dt={cfg.delta_t}, one passive scalar, FP64 RK3/CN",
                   "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz, "n_th": cfg.n_th, "precision": "fp64",
                   "substeps_per_step": 3, "stage_model_bytes_per_substep": model},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "batched": batched,
        "les": les,
        "zdft": zd,
        "autotune": tuned if args.autotune == "on" else None,
        "graphs": graphs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
