"""Benchmark of the RHS + SSP-RK3 path on one MI355X.

Workload: the Taylor-Green vortex on a 256^3 periodic block over [0, 2pi]^3, characteristic-wise
WENO-5 (Shu weights, Lax-Friedrichs split) + CD4 viscous terms with the Sutherland law, FP64,
SSP-RK3 at CFL 0.5 (--viscosity constant runs the reference's constant-mu path instead).
A step is one full SSP-RK3 step (3 RHS evaluations + CFL reduction + stage updates) with the state
resident in HBM.  Metric: Gcell-RK3-steps/s = nx*ny*nz*steps / wall time of the timed steps.

Prints one JSON line (see the keys at the end of main()).
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic work per unit (one cell advanced one SSP-RK3 step, characteristic WENO-5 + CD4,
# Shu weights; add/sub/mul = 1, fma = 2, div = 1, sqrt = 1): 12,050 FP64 flop.  Per face of one
# characteristic sweep: 1,190 flop; per face of one component-wise sweep: 725 flop.
FLOP_PER_CELL_STEP = {"weno5cw": 12050.0, "weno5": 7860.0}
FLOP_PER_FACE = {"weno5cw": 1190.0, "weno5": 725.0}
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector = matrix), dense
# planes the residual + stage-update kernel moves per launch, each once: 31 read (fx, fy, fz, Un,
# Ucur: 5 each; W: 4; dW/dx of u, v: 2; it forms the other derivatives itself) and 10 written
# (Uout, Wout); FP64 planes, except the face fluxes of the FP32 build (FP32 planes)
RES_PLANES = 41
# with the z sweep fused into it, the residual kernel reads U and W of the z stencils instead of
# fz: 31 read (fx, fy, Un, Ucur: 5 each; W: 5; dW/dx of u, v: 2; the first face of each row: a
# plane of 1/n) and 10 written
RES_PLANES_FUSED_Z = 41
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, dense


def cpu_baseline_child(n, steps):
    """Times the reference (oracle/_ref/libnsref_fast.so: the reference's own translation units,
    its optimisation flags) or, when that was not built, the oracle, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    C = importlib.import_module("3dnssolver_amd.cases")
    if pyoracle.reference_available("fast"):
        lib, kind = pyoracle.Reference("fast"), "reference"
    else:
        lib, kind = pyoracle.Oracle(), "port"
    warm_case, warm_W = C.config2(32)
    lib.step(warm_case, warm_W, 1)
    case, W = C.config2(n)
    t0 = time.perf_counter()
    lib.step(case, W, steps)
    el = time.perf_counter() - t0
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    print(json.dumps({"value": n ** 3 * steps / el / 1e9, "unit": "Gcell-RK3-steps/s",
                      "cores": lib.num_threads(), "kind": kind, "cpu_model": model,
                      "logical_cpus": os.cpu_count(),
                      "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                             "OMP_PLACES")},
                      "sample": f"{steps} SSP-RK3 step(s) of the same workload at {n}^3 "
                                f"({el:.1f} s wall, OpenMP threads = cores), constant mu: the "
                                f"reference's code path never calls its Sutherland law"}))


def run_cpu_baseline(n, steps, timeout):
    try:
        # one OpenMP thread per core the job may use, spread over the cores
        env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
                            "--cpu-n", str(n), "--cpu-steps", str(steps)],
                           capture_output=True, text=True, timeout=timeout, env=env)
        lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode == 0 and lines:
            return json.loads(lines[-1])
        sys.stderr.write(r.stderr[-2000:])
    except subprocess.TimeoutExpired:
        sys.stderr.write("cpu baseline timed out\n")
    return None


def precision_workload(args):
    """Precision form (FP32 reconstruction + viscous fluxes, FP64 accumulation) on decaying
    isotropic turbulence at n^3: the FP64 path and the FP32 path advance the same seeded state
    the same number of steps; reports both rates and the L2 / Linf errors of density and total
    energy of the FP32 run against the FP64 run."""
    import numpy as np
    ns = importlib.import_module("3dnssolver_amd")
    C = importlib.import_module("3dnssolver_amd.cases")
    n = args.n
    print(f"[bench] precision form {n}^3: generating the isotropic field", file=sys.stderr, flush=True)
    case, W0 = C.config3(n)
    out = {}
    states = {}
    for math in ("exact", "fp32"):
        s = ns.Solver(case, math=math)
        s.set_state(W0, "primitive")
        s.step(args.warmup)
        print(f"[bench] {math}: {args.steps} steps", file=sys.stderr, flush=True)
        el, done = 0.0, 0
        while done < args.steps:  # chunks of 100 steps, one progress line each
            k = min(100, args.steps - done)
            t0 = time.perf_counter()
            s.step(k)
            el += time.perf_counter() - t0
            done += k
            print(f"[bench] {math}: {done}/{args.steps} steps", file=sys.stderr, flush=True)
        out[math] = n ** 3 * args.steps / el / 1e9
        U = s.get_state("conservative")[3:-3, 3:-3, 3:-3]
        states[math] = (U[..., 0].copy(), U[..., 1].copy())
        _, _, flags, rc = s.reduce()
        out[math + "_health"] = {"status": rc, "flags": flags}
        # per-kernel times of one more step (HIP events on the context's stream)
        s.timing(True)
        s.step(1)
        out[math + "_timers"] = s.timers()
        s.timing(False)
        s.close()
        del U
    errs = {}
    for q, name in ((0, "rho"), (1, "rhoE")):
        a, b = states["fp32"][q], states["exact"][q]
        d = a - b
        errs[name] = {"L2_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(b * b))),
                      "Linf_rel": float(np.abs(d).max() / np.abs(b).max()),
                      "L2_abs": float(np.sqrt(np.mean(d * d))), "Linf_abs": float(np.abs(d).max())}
    # roofline of the FP32 face kernel (the dominant sweep) and of the HBM-bound residual
    tm = out["fp32_timers"]
    flux = {k: v for k, v in tm.items() if k.startswith("face_flux")}
    dom = max(flux, key=lambda k: flux[k][0])
    dom_ms = flux[dom][0] / flux[dom][1]
    flop = FLOP_PER_FACE["weno5cw"] * (n + 1) * n * n
    fp32_tf = flop / (dom_ms * 1e-3) / 1e12
    res_ms = tm["residual_update"][0] / tm["residual_update"][1]
    # the FP32 build's residual reads its 15 face-flux planes in single precision
    res_gbs = ((RES_PLANES - 15) * 8.0 + 15 * 4.0) * n ** 3 / (res_ms * 1e-3) / 1e9
    fp64_dom = max((k for k in out["exact_timers"] if k.startswith("face_flux")),
                   key=lambda k: out["exact_timers"][k][0])
    fp64_ms = out["exact_timers"][fp64_dom][0] / out["exact_timers"][fp64_dom][1]
    print(json.dumps({
        "metric": "precision form: Gcell-RK3-steps/s, FP32 reconstruction vs FP64",
        "value": round(out["fp32"], 6), "unit": "Gcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"decaying isotropic turbulence {n}^3, [0,2pi]^3, E(k)~k^4 exp(-2(k/4)^2), "
                               f"Mt 0.5, Re_lambda 100, WENO-5cw + CD4, SSP-RK3 CFL {case['cfl']}",
                   "grid": [n, n, n]},
        "fp64_value": round(out["exact"], 6), "speedup_vs_fp64": round(out["fp32"] / out["exact"], 4),
        "errors_vs_fp64": errs,
        "health": {"fp32": out["fp32_health"], "fp64": out["exact_health"]},
        "roofline": {"bound": "valu", "compute": "FP32 vector ALU (no packed FP32 instructions; "
                                                 "FP32 ones issue faster than FP64 ones)",
                     "kernel": dom, "achieved": round(fp32_tf, 4), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(fp32_tf / FP32_PEAK_TFLOPS, 5),
                     "traffic": None, "flop_per_launch": flop, "avg_launch_ms": round(dom_ms, 5),
                     "fp64_same_kernel_ms": round(fp64_ms, 5)},
        "roofline_residual": {"bound": "hbm", "kernel": "residual_update",
                              "achieved": round(res_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(res_gbs / HBM_PEAK_GBS, 4), "avg_launch_ms": round(res_ms, 5)},
        "kernel_ms_per_step": {m: {k: round(v[0], 4) for k, v in out[m + "_timers"].items()}
                               for m in ("exact", "fp32")},
    }))


def ensemble_workload(args):
    """Batched form: 64 independent periodic shear layers of uneven sizes advanced by one
    hipGraph-captured SSP-RK3 step over all blocks (FP32 reconstruction, FP64 accumulation),
    tiles from the autotuner.  Reports the aggregate rate, the rate of the same blocks stepped one
    after another through the single-block path, and per-instance relative L2 errors of density
    and total energy against the plain FP64 method run on each block alone."""
    import numpy as np
    ns = importlib.import_module("3dnssolver_amd")
    C = importlib.import_module("3dnssolver_amd.cases")
    members = C.config5(n_blocks=args.blocks)
    cells = sum(c["nx"] * c["ny"] * c["nz"] for c, _, _ in members)
    print(f"[bench] ensemble of {len(members)} blocks, {cells / 1e6:.2f} Mcells, math={args.math}",
          file=sys.stderr, flush=True)
    params = None
    if args.weights == "learned":
        params = importlib.import_module("3dnssolver_amd.learned").load()
    solvers = []
    for case, W, _ in members:
        s = ns.Solver(dict(case, weights=args.weights), math=args.math)
        if params is not None:
            s.set_mlp(params)
        s.set_state(W, "primitive")
        solvers.append(s)
    ens = ns.Ensemble(solvers)
    ms_default = ms_tuned = None
    if not args.no_tune:
        ms_default, ms_tuned = ens.tune(steps=3)
        print(f"[bench] autotuner: {ms_default:.3f} -> {ms_tuned:.3f} ms/step", file=sys.stderr,
              flush=True)
    tiles = ens.tiles
    ens.step(args.warmup)
    t0 = time.perf_counter()
    ens.step(args.steps)
    el = time.perf_counter() - t0
    value = cells * args.steps / el / 1e9
    nog = 3
    t0 = time.perf_counter()
    ens.step(nog, graph=False)
    ms_nograph = (time.perf_counter() - t0) / nog * 1e3
    done = args.warmup + args.steps + nog
    # the same blocks one after another through the single-block path, same math (timed on
    # copies of the initial states)
    t_solo = 0.0
    for case, W, _ in members:
        s = ns.Solver(dict(case, weights=args.weights), math=args.math)
        if params is not None:
            s.set_mlp(params)
        s.set_state(W, "primitive")
        s.step(1)
        t0 = time.perf_counter()
        s.step(nog)
        t_solo += time.perf_counter() - t0
        s.close()
    ms_solo = t_solo / nog * 1e3
    # per-instance error against the plain FP64 method (EXACT math) on the block alone
    print(f"[bench] FP64 reference runs ({done} steps per block)", file=sys.stderr, flush=True)
    l2 = {"rho": [], "rhoE": []}
    health = []
    for (case, W, _), s in zip(members, solvers):
        U = s.get_state("conservative")[3:-3, 3:-3, 3:-3]
        health.append(s.reduce()[3])
        r = ns.Solver(case, math="exact")
        r.set_state(W, "primitive")
        r.step(done)
        Ur = r.get_state("conservative")[3:-3, 3:-3, 3:-3]
        r.close()
        for q, name in ((0, "rho"), (1, "rhoE")):
            d = U[..., q] - Ur[..., q]
            l2[name].append(float(np.sqrt(np.sum(d * d) / np.sum(Ur[..., q] ** 2))))
    ens.close()
    for s in solvers:
        s.close()
    print(json.dumps({
        "metric": "batched form: aggregate Gcell-RK3-steps/s over the ensemble",
        "value": round(value, 6), "unit": "Gcell-RK3-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"exact": "f64", "fast": "f64", "fp32": "f32+f64"}[args.math], "data": "synthetic",
        "config": {"workload": f"{len(members)} periodic double shear layers on [0,1]^3, edges from "
                               f"{{32,48,64,96,128}} per axis, Mach U[0.2,0.6], Re U[500,5000], "
                               f"amplitude U[0.01,0.1], c0 = 0.9*3/(1+M), WENO-5 (Shu, LF) + CD4, "
                               f"SSP-RK3 CFL {members[0][0]['cfl']}",
                   "cells": cells, "math": args.math, "weights": args.weights, "graph": True},
        "autotuner": {"ms_default_tiles": ms_default, "ms_tuned": ms_tuned,
                      "tiles_kw_flux_kw_cell_march": tiles.tolist()},
        "ms_per_step_without_graph": round(ms_nograph, 4),
        "ms_per_step_blocks_one_at_a_time": round(ms_solo, 4),
        "l2_rel_vs_fp64_plain": {k: {"max": max(v), "median": float(np.median(v))}
                                 for k, v in l2.items()},
        "l2_rel_per_instance": l2, "steps_compared": done,
        "health": {"nonzero": [h for h in health if h != 0]},
    }))


def learned_workload(args):
    """Learned form: the double shear layer at n^3 (component-wise WENO-5 + CD4).  Times SSP-RK3
    steps of the plain FP64 method (Shu weights, EXACT math), of FP32 reconstruction with the Shu
    weights and with the learned weights, and of FP64 with the learned weights; then, for the
    shear layer and the Sod-like planar jump, the relative L2 errors of density and energy of
    the plain FP64 run and of the FP32 + learned run against a 2n^3 FP64 run at the same time
    (prescribed dt)."""
    import numpy as np
    ns = importlib.import_module("3dnssolver_amd")
    C = importlib.import_module("3dnssolver_amd.cases")
    lr = importlib.import_module("3dnssolver_amd.learned")
    params = lr.load()
    n = args.n
    case, W = C.config4(n, "shear")
    rates = {}
    for math, weights in (("exact", "shu"), ("fp32", "shu"), ("fp32", "learned"),
                          ("exact", "learned")):
        s = ns.Solver(dict(case, weights=weights), math=math)
        if weights == "learned":
            s.set_mlp(params)
        s.set_state(W, "primitive")
        s.step(args.warmup)
        print(f"[bench] {math}/{weights}: {args.steps} steps at {n}^3", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        s.step(args.steps)
        el = time.perf_counter() - t0
        rates[f"{math}_{weights}"] = {"value": round(n ** 3 * args.steps / el / 1e9, 6),
                                      "ms_per_step": round(el / args.steps * 1e3, 4),
                                      "health": s.reduce()[3]}
        s.close()
    errors = {}
    ne = args.eval_n
    for prob in ("shear", "jump") if args.eval_steps > 0 else ():  # --eval-steps 0: rates only
        print(f"[bench] accuracy: {prob} {ne}^3 vs {2 * ne}^3 FP64", file=sys.stderr, flush=True)
        c, w = C.config4(ne, prob)
        fc, fw = C.config4(2 * ne, prob)
        dt, t_end, ref = lr.reference(c, w, fc, fw, args.eval_steps)
        e = {"t_end": t_end, "steps": args.eval_steps, "n": ne}
        for math, weights in (("exact", "shu"), ("fp32", "shu"), ("fp32", "learned")):
            e[f"{math}_{weights}"] = lr.coarse_error(c, w, dt, args.eval_steps, ref, math,
                                                     weights, params)
        errors[prob] = e
    base = rates["exact_shu"]["value"]
    print(json.dumps({
        "metric": "learned form: FP32 + learned weights vs plain FP64, Gcell-RK3-steps/s and L2 "
                  "vs a 2x-resolution FP64 run",
        "value": rates["fp32_learned"]["value"], "unit": "Gcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
        "dtype": "f32+bf16 network+f64", "data": "synthetic",
        "config": {"workload": f"double shear layer {n}^3 on [0,1]^3, Mach 0.3, Re 1000, "
                               f"WENO-5 component-wise (LF) + CD4, SSP-RK3 CFL 0.5; accuracy on "
                               f"the shear layer and a Sod-like planar jump (8:1, 10:1) at "
                               f"{ne}^3 vs {2 * ne}^3", "grid": [n, n, n]},
        "speedup_vs_fp64_plain": round(rates["fp32_learned"]["value"] / base, 4),
        "rates": rates, "l2_vs_2x_fp64": errors,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["tgv", "precision", "ensemble", "learned"],
                    default="tgv",
                    help="tgv: the headline run; precision: FP32-reconstruction form; "
                         "ensemble: batched form; learned: learned-weights form")
    ap.add_argument("--weights", choices=["shu", "learned"], default="shu",
                    help="WENO weights of the ensemble members")
    ap.add_argument("--eval-n", type=int, default=64)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--math", choices=["exact", "fast", "fp32"], default="exact")
    ap.add_argument("--conv", choices=["weno5cw", "weno5"], default="weno5cw")
    ap.add_argument("--viscosity", choices=["constant", "sutherland"], default="sutherland",
                    help="the Taylor-Green workload's Sutherland law, or constant mu (the only "
                         "law the reference's code path calls)")
    ap.add_argument("--alpha-mode", choices=["const3", "local_max"], default="const3",
                    help="Lax-Friedrichs constant: the reference's 3.0, or per-face eigenvalue")
    ap.add_argument("--physics", choices=["reference", "corrected_convection", "corrected_viscous",
                                          "corrected"], default="reference")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-baseline-child", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32 and learned sub-keys of the headline line")
    ap.add_argument("--learned-steps", type=int, default=10)
    args = ap.parse_args()

    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_n, args.cpu_steps)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "precision":
        precision_workload(args)
        return
    if args.workload == "ensemble":
        ensemble_workload(args)
        return
    if args.workload == "learned":
        learned_workload(args)
        return

    ns = importlib.import_module("3dnssolver_amd")
    C = importlib.import_module("3dnssolver_amd.cases")
    case, W0 = C.config2(args.n, viscosity=args.viscosity)
    case["conv"] = args.conv
    case["alpha_mode"], case["physics"] = args.alpha_mode, args.physics
    n = args.n
    cells = n ** 3
    print(f"[bench] {n}^3 {args.conv} cd4 math={args.math}: setup", file=sys.stderr, flush=True)
    s = ns.Solver(case, math=args.math)
    t_up = time.perf_counter()
    s.set_state(W0, "primitive")
    t_up = time.perf_counter() - t_up  # host array -> device (PCIe) + layout conversion
    if args.warmup:
        s.step(args.warmup)
    print(f"[bench] warmup done, timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    _, dt = s.step(args.steps)          # returns after hipStreamSynchronize
    el = time.perf_counter() - t0
    sums, dtmin, flags, rc = s.reduce()
    value = cells * args.steps / el / 1e9
    ms_per_step = el / args.steps * 1e3
    t_down = time.perf_counter()
    U_fp64 = s.get_state("conservative")[3:-3, 3:-3, 3:-3]  # device -> host array
    t_down = time.perf_counter() - t_down
    pcie_value = cells * args.steps / (el + t_up + t_down) / 1e9

    # per-kernel durations with HIP events on the context's stream (separate pass, one stream)
    s.timing(True)
    s.step(args.timing_steps)
    timers = s.timers()
    s.timing(False)
    s.close()
    kernels = kernel_rooflines(timers, args, n)
    step_tflops = value * 1e9 * FLOP_PER_CELL_STEP[args.conv] / 1e12
    traffic = step_traffic(timers, args, n)

    forms = None
    if not args.no_forms and args.math == "exact" and args.conv == "weno5cw":
        forms = {"fp32": fp32_form(ns, case, W0, args, value, U_fp64),
                 "fast": fp32_form(ns, case, W0, args, value, U_fp64, math="fast")}
        del U_fp64
        forms["learned"] = learned_form(ns, C, args)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = run_cpu_baseline(args.cpu_n, args.cpu_steps, timeout=240)

    out = {
        "metric": "Gcell-RK3-steps/s at 256^3 + % binding roofline; learned/FP32 vs plain FP64 at L2",
        "value": round(value, 6),
        "unit": "Gcell-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"Taylor-Green vortex {n}^3 periodic, [0,2pi]^3, Mach 0.1, Re 1600, "
                        f"{'characteristic' if args.conv == 'weno5cw' else 'component-wise'} "
                        f"WENO-5 (Shu, LF split) + CD4 viscous ({args.viscosity} mu), SSP-RK3, "
                        f"CFL 0.5",
            "grid": [n, n, n], "math": args.math, "viscosity": args.viscosity,
            "alpha_mode": args.alpha_mode, "physics": args.physics,
            "state_resident_in_hbm": True,
        },
        # the whole SSP-RK3 step against the FP64 vector peak: the algorithmic flop of one
        # cell-step (12,050 characteristic / 7,860 component-wise) x cells x steps / wall time
        "roofline": {
            "bound": "valu", "compute": "FP64 vector ALU (MI355X FP64 peak 78.6 TFLOP/s)",
            "kernel": "whole SSP-RK3 step (all kernels)",
            "achieved": round(step_tflops, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(step_tflops / FP64_PEAK_TFLOPS, 5),
            "flop_per_cell_step": FLOP_PER_CELL_STEP[args.conv],
            "avg_launch_ms": round(ms_per_step, 4),
            "traffic": traffic["bytes_per_step"],
            "traffic_bytes_per_cell_step": traffic["bytes_per_cell_step"],
            "ideal_bytes_per_cell_step": 320,
            "traffic_source": traffic["source"],
            "kernels": kernels,
        },
        "kernel_ms_per_step": {k: round(v[0] / args.timing_steps, 4) for k, v in timers.items()},
        "health": {"status": rc, "flags": flags, "dt_last": dt},
        "pcie_inclusive": {"value": round(pcie_value, 6), "upload_ms": round(t_up * 1e3, 2),
                           "download_ms": round(t_down * 1e3, 2),
                           "note": "the timed steps plus one upload of the host state before and "
                                   "one download after"},
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def kernel_rooflines(timers, args, n):
    """Per-launch rooflines of the dominant face-flux kernel (FP64 VALU) and of the residual +
    stage-update kernel (HBM), from the HIP-event timing pass."""
    out = {}
    flux = {k: v for k, v in timers.items() if k.startswith("face_flux") and v[1]}
    if flux:
        dom = max(flux, key=lambda k: flux[k][0] / flux[k][1])
        ms = flux[dom][0] / flux[dom][1]
        flop = FLOP_PER_FACE[args.conv] * (n + 1) * n * n
        tf = flop / (ms * 1e-3) / 1e12
        out["face_flux"] = {"bound": "valu", "kernel": dom, "achieved": round(tf, 4),
                            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(tf / FP64_PEAK_TFLOPS, 5), "flop_per_launch": flop,
                            "avg_launch_ms": round(ms, 5)}
    res = timers.get("residual_update")
    if res and res[1]:
        ms = res[0] / res[1]
        planes = RES_PLANES_FUSED_Z if timers.get("face_flux_z", (0, 0))[1] == 0 else RES_PLANES
        # the FP32 build stores the 15 face-flux planes it reads in single precision
        flux_bytes = 4.0 if args.math == "fp32" else 8.0
        b = ((planes - 15) * 8.0 + 15 * flux_bytes) * n ** 3
        gbs = b / (ms * 1e-3) / 1e9
        out["residual_update"] = {"bound": "hbm", "kernel": "residual_update",
                                  "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": b,
                                  "avg_launch_ms": round(ms, 5)}
    return out


def step_traffic(timers, args, n):
    """HBM bytes of one step: the launches of one step times each kernel's bytes per launch from
    the committed counter summary (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this workload,
    calibrated on a copy kernel): a lookup, not a measurement of this run."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    none = {"bytes_per_step": None, "bytes_per_cell_step": None, "source": None}
    if not os.path.exists(path):
        return none
    tab = json.load(open(path))
    key = f"{args.math}_{n}_{args.viscosity}"
    per = tab.get("per_step", {}).get(key)
    if per is None:
        return none
    return {"bytes_per_step": per, "bytes_per_cell_step": round(per / n ** 3, 1),
            "source": f"committed lookup profiles/traffic.json[per_step][{key}] (rocprofv3 --pmc "
                      f"FETCH_SIZE + WRITE_SIZE per kernel, copy-kernel calibrated), not measured "
                      f"in this run"}


def _l2(a, b):
    import numpy as np
    d = a - b
    return float(np.sqrt(np.sum(d * d) / np.sum(b * b)))


def fp32_form(ns, case, W0, args, fp64_value, U_fp64, math="fp32"):
    """Precision form on the headline workload: the same steps from the same state in FP32
    reconstruction + FP64 accumulation; rate, speed-up and the relative L2 / Linf of density and
    total energy against the FP64 (EXACT) state after the same steps.  math="fast": the same for
    FP64 with contraction (fused multiply-add within expressions) and hoisted reciprocals."""
    import numpy as np
    n = args.n
    s = ns.Solver(case, math=math)
    s.set_state(W0, "primitive")
    if args.warmup:
        s.step(args.warmup)
    t0 = time.perf_counter()
    s.step(args.steps)
    el = time.perf_counter() - t0
    U = s.get_state("conservative")[3:-3, 3:-3, 3:-3]
    rc = s.reduce()[3]
    s.close()
    v = n ** 3 * args.steps / el / 1e9
    err = {}
    for q, name in ((0, "rho"), (1, "rhoE")):
        a, b = U[..., q], U_fp64[..., q]
        err[name] = {"L2_rel": _l2(a, b), "Linf_rel": float(np.abs(a - b).max() / np.abs(b).max())}
    print(f"[bench] {math} form: {v:.3f} Gcell-steps/s", file=sys.stderr, flush=True)
    what = ("the headline workload, FP32 reconstruction and viscous fluxes, FP64 state, stage "
            "accumulation and reductions" if math == "fp32" else
            "the headline workload in FP64 with fused multiply-add within expressions and hoisted "
            "reciprocals (round-off of the reference's order, not bitwise)")
    tr = step_traffic(None, argparse.Namespace(math=math, viscosity=args.viscosity), n)
    return {"workload": what,
            "value": round(v, 6), "ms_per_step": round(el / args.steps * 1e3, 4),
            "speedup_vs_fp64": round(v / fp64_value, 4),
            "errors_vs_fp64": err, "steps_compared": args.warmup + args.steps, "health": rc,
            "traffic": tr["bytes_per_step"], "traffic_bytes_per_cell_step": tr["bytes_per_cell_step"],
            "traffic_source": tr["source"]}


def learned_form(ns, C, args):
    """Learned form: the 256^3 shear layer (component-wise WENO-5 + CD4) stepped with the plain
    FP64 method (Shu weights, EXACT) and with FP32 reconstruction + the learned weights; rates
    and speed-up, and the relative L2 errors of both against a 2x-resolution FP64 run on the
    shear layer and the Sod-like jump (args.eval_n^3 against (2 eval_n)^3, prescribed dt)."""
    lr = importlib.import_module("3dnssolver_amd.learned")
    params = lr.load()
    n = args.n
    case, W = C.config4(n, "shear")
    rates = {}
    for math, weights in (("exact", "shu"), ("fp32", "learned")):
        s = ns.Solver(dict(case, weights=weights), math=math)
        if weights == "learned":
            s.set_mlp(params)
        s.set_state(W, "primitive")
        s.step(2)
        t0 = time.perf_counter()
        s.step(args.learned_steps)
        el = time.perf_counter() - t0
        rates[f"{math}_{weights}"] = {"value": round(n ** 3 * args.learned_steps / el / 1e9, 6),
                                      "ms_per_step": round(el / args.learned_steps * 1e3, 4),
                                      "health": s.reduce()[3]}
        s.close()
    errors = {}
    ne = args.eval_n
    for prob in ("shear", "jump"):
        c, w = C.config4(ne, prob)
        fc, fw = C.config4(2 * ne, prob)
        dtp, t_end, ref = lr.reference(c, w, fc, fw, args.eval_steps)
        e = {"t_end": t_end, "steps": args.eval_steps}
        for math, weights in (("exact", "shu"), ("fp32", "learned")):
            e[f"{math}_{weights}"] = lr.coarse_error(c, w, dtp, args.eval_steps, ref, math,
                                                     weights, params)
        errors[prob] = e
    sp = rates["fp32_learned"]["value"] / rates["exact_shu"]["value"]
    print(f"[bench] learned form: {sp:.3f}x plain FP64", file=sys.stderr, flush=True)
    return {"workload": f"double shear layer {n}^3 on [0,1]^3, Mach 0.3, Re 1000, component-wise "
                        f"WENO-5 (LF) + CD4, SSP-RK3 CFL 0.5; accuracy on the shear layer and the "
                        f"Sod-like planar jump (8:1, 10:1) at {ne}^3 against a {2 * ne}^3 FP64 run",
            "rates": rates, "speedup_vs_fp64_plain": round(sp, 4),
            "l2_rel_vs_2x_fp64": errors}


if __name__ == "__main__":
    main()
