"""Benchmark of the per-step phase-space update on one MI355X.

Default workload: the velocity-adaptive cylinder (example/cylinder/cylinder.jl at Ma = 5):
Circle(Maxwellian, r = 0.5) immersed in [-8,8]^2 on 448 x 448 leaves (200,704), Kn = 0.1,
omega = 0.81, a developed bow-shock state, per-cell adapted velocity grids on a 20 x 20 base
on [-15,15]^2 with 2 levels (568-4000 cells per grid: the reference's initial grid of each
cell's Maxwellian, refined toward 4000 cells with the local disturbance), CAIDVM flux +
CAIDVM_Marching with the immersed boundary, in the FP32 form (f stored and advected in FP32,
compensated moment sums and the Maxwellian parameters in FP64), each step replayed as one HIP
graph whose independent launches run as parallel branches on four streams, with the launch
options first tuned on the workload (coordinate descent; the choice is reported). The FP64
form runs beside it on the same workload. One step = slope + IB + flux + relaxation over all
fluid phase-space cells.

Prints one JSON line (rank 0). The CPU baseline is the oracle (a C++ FP64 restatement of the
reference algorithm, OpenMP over cells and faces) timed on this host on whole steps of the
same workload, with all host threads and with one.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kitamr.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

HBM_PEAK_GBS = 8000.0


def alg_bytes(s_f, adaptive=False, dim=2, ndf=2):
    """Algorithmic bytes per phase-space update of the reference's 3-phase dataflow, with
    s_f bytes per stored f value (8: FP64, 4: FP32 form): slope R df (+ R level) + W sdf;
    flux R df, R sdf (+ R u, R weight, R level), W flux; iterate R df, R flux (+ R u,
    R weight), W df. The geometry terms (8-byte u components and weight, 1-byte level)
    count only for per-cell adapted velocity grids; a shared grid is read once.
    2D2F FP64: 48, 112, 160 (adaptive 210); 2D2F FP32 form adaptive: 25, 105, 130;
    3D1F FP64: 32, 64, 96 (adaptive 162)."""
    g_slope, g_flux, g_iter = ((1, dim * 8 + 8 + 1, dim * 8 + 8) if adaptive else (0, 0, 0))
    b_slope = ndf * s_f * (1 + dim) + g_slope
    b_flux_iter = ndf * s_f * (1 + dim + 1) + g_flux + ndf * s_f * 3 + g_iter
    return b_slope, b_flux_iter, b_slope + b_flux_iter


WORKLOADS = {
    "cylinder2d-vsamr": "2D cylinder (velocity-adaptive), Circle(Maxwellian, r=0.5) immersed in "
                        "[-8,8]^2, 448x448 leaves (200704), Mach-5 inflow, Kn=0.1, omega=0.81, "
                        "developed bow-shock state, per-cell adapted velocity grids (20x20 base "
                        "on [-15,15]^2, 2 levels, 568-4000 cells), CAIDVM + CAIDVM_Marching with "
                        "the immersed boundary",
    "riemann2d": "2D Riemann, 64x64 base + 3 static PS levels (49312 leaves), uniform 48x48 "
                 "velocity grid, CAIDVM + CAIDVM_Marching",
    "riemann2d-c1": "2D Riemann, 64x64 uniform leaves, uniform 28x28 velocity grid on [-5,5]^2, "
                    "Kn=1e-2, omega=0.5, CAIDVM + CAIDVM_Marching",
    "riemann2d-vsamr": "2D Riemann, 64x64 base + 3 static PS levels (49312 leaves), per-cell "
                       "adapted velocity grids (20x20 base on [-15,15]^2, 2 levels, 1567-1651 "
                       "cells each), CAIDVM + CAIDVM_Marching",
    "cylinder2d": "2D cylinder, Circle(Maxwellian, r=0.5) immersed in [-8,8]^2, 448x448 "
                  "cells, Mach-5 inflow, Kn=0.1, uniform 40x40 velocity grid on [-15,15]^2, "
                  "CAIDVM + CAIDVM_Marching with update_solid_cell! / update_solid_neighbor!",
    "sphere3d-vsamr": "3D1F sphere (velocity-adaptive), Sphere(Maxwellian, r=0.5) immersed in "
                      "[-6,6]^3, 48^3 leaves (110592), Mach-3 inflow, Kn=0.05, omega=0.81, "
                      "developed bow-shock state, per-cell adapted velocity grids (16^3 base on "
                      "[-10,10]^3, 1 level, 4096-30000 cells), CAIDVM + CAIDVM_Marching with "
                      "the immersed boundary",
}


def make_case(workload, precision):
    from kitamr_hip import cases
    if workload == "cylinder2d-vsamr":
        return cases.cylinder_c3(n=448, precision=precision)
    if workload == "riemann2d-vsamr":
        return cases.riemann_2d(n=64, nv=20, vmax=15.0, ps_levels=3, vs_levels=2,
                                vs_variants=6, precision=precision)
    if workload == "sphere3d-vsamr":
        return cases.sphere_c4(n=48, precision=precision)
    if workload == "cylinder2d":
        return cases.cylinder_2d(n=448, nv=40, vmax=15.0, L=8.0, radius=0.5, Ma=5.0, Kn=0.1,
                                 omega=0.81, precision=precision)
    if workload == "riemann2d-c1":
        return cases.riemann_2d(n=64, nv=28, precision=precision)
    return cases.riemann_2d(n=64, nv=48, ps_levels=3, precision=precision)


def with_precision(case, precision):
    """The same case (topology and initial state) in another storage form."""
    return dataclasses.replace(case, config=dataclasses.replace(case.config,
                                                                precision=precision))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def torch_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


PROFILES = {
    # (workload, precision) -> committed rocprofv3 summary with the counter traffic of the
    # dominant kernel (k_flux_relax) on that workload
    ("cylinder2d-vsamr", "fp32f"): ROOT / "profiles" / "k_flux_relax_cylinder2d_c3_200k_fp32.md",
    ("riemann2d", "fp64"): ROOT / "profiles" / "k_flux_relax_riemann2d_49k_48x48.md",
}


def profiled_traffic(workload, precision):
    """HBM bytes per k_flux_relax launch on this workload, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (counters scaled by the copy-kernel calibration)."""
    path = PROFILES.get((workload, precision))
    if path is None:
        return None
    try:
        for line in path.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch k_flux_relax = "):
                return float(line.rsplit("=", 1)[1])
    except OSError:
        pass
    return None


def serving_stream(n_req=16, seed=20261202, rate=2.0, budget_gb=200.0, slots=4, max_steps=100):
    """A stream of the serving workload: requests drawn from its distributions (serve.c5_specs:
    Riemann or cylinder, Kn log-uniform in [1e-3, 1], Ma in [0.5, 5], 4k-200k cells, 16^2-64^2
    velocity grids, FP64 or the FP32 form), Poisson arrivals; each runs to the residual
    tolerance or max_steps (reduced from 2000 to keep the run short). Four slots (one context
    each), as the GPU has four hardware queues per process. Reports p50/p99 request latency
    (arrival to result) and the aggregate updates/s."""
    from kitamr_hip import serve
    specs = serve.c5_specs(n_req, seed)
    arrivals = serve.poisson_arrivals(n_req, rate, seed)
    reqs = []
    t0 = time.perf_counter()
    for k, sp in enumerate(specs):
        reqs.append(serve.Request(rid=k, case=serve.c5_case(sp), arrival=float(arrivals[k]),
                                  max_steps=max_steps, tag=serve.c5_tag(sp)))
    log(f"[bench] serving stream of {n_req} requests built in {time.perf_counter() - t0:.1f}s")
    srv = serve.Server(budget_bytes=int(budget_gb * 1e9), max_slots=slots)
    rep = srv.run(reqs, keep_df=False)
    rep["workload"] = (f"{n_req} requests, Poisson arrivals at {rate}/s, drawn from the serving "
                       "distributions: 2D Riemann (outflow) or cylinder (immersed circle, Mach "
                       "0.5-5 inflow), Kn log-uniform in [1e-3, 1], 4k-200k cells, 16^2-64^2 "
                       f"velocity grids, FP64 or the FP32 form, each to the residual tolerance or "
                       f"{max_steps} steps (step budget reduced from 2000 to keep the run short)")
    rep["slots"] = slots
    rep["phase_cells_per_request"] = [int(r.case.phase_cells) for r in reqs]
    return rep


def cpu_baseline(case, max_steps=3, min_seconds=10.0, one_core_steps=1):
    """The oracle (C++ FP64 restatement of the reference's step) on this host: whole steps of
    the same workload with every host thread, then with one thread (the reference runs one
    single-threaded rank per core, src/KitAMR.jl:44-46)."""
    import numpy as np
    import oracle as O
    from kitamr_hip.solver import kah_params
    c64 = with_precision(case, "fp64")
    orc = O.Oracle(kah_params(c64.config), c64.topo)
    orc.set_state(c64.w, c64.prim, c64.df)
    threads = O.set_threads(0)
    times = []
    t_all = time.perf_counter()
    step = 0
    for _ in range(max_steps):
        t0 = time.perf_counter()
        orc.step(c64.dt_xi, step)
        step += 1
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all >= min_seconds:
            break
    med = float(np.median(times))
    O.set_threads(1)
    t1 = []
    for _ in range(one_core_steps):
        t0 = time.perf_counter()
        orc.step(c64.dt_xi, step)
        step += 1
        t1.append(time.perf_counter() - t0)
    O.set_threads(0)
    orc.close()
    med1 = float(np.median(t1))
    return {"value": c64.phase_cells / med, "unit": "updates/s", "cores": int(threads),
            "kind": "port",
            "sample": f"{len(times)} full steps of the same workload on the oracle (C++ FP64 "
                      f"restatement, OpenMP {threads} threads), median step {med:.3f} s; then "
                      f"{len(t1)} step(s) on 1 thread",
            "one_core": {"value": c64.phase_cells / med1, "unit": "updates/s", "cores": 1,
                         "step_s": med1}}


def timed_run(case, options, steps, warmup, graphs, copy=False, log=None):
    """Fused steps of one case on a fresh context. Returns a dict: wall time of the timed
    steps (graph replay when graphs), the per-kernel average launch times from HIP events on
    the library's stream over the same number of eager steps, finiteness, copy bandwidth."""
    import numpy as np
    from kitamr_hip import abi
    from kitamr_hip.solver import kah_params
    dev = abi.HipContext(kah_params(case.config))
    for k, v in options.items():
        dev.set_option(k, v)
    dev.set_option("graphs", 1 if graphs else 0)
    t0 = time.perf_counter()
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    dev.sync()
    t_setup = time.perf_counter() - t0
    dt = case.dt_xi
    step = 0
    for _ in range(warmup):
        dev.step(dt, step, residual=False)
        step += 1
    dev.sync()
    log(f"[bench] setup {t_setup:.1f}s, warmup {warmup} steps done; device bytes "
        f"{dev.device_bytes() / 1e9:.2f} GB")
    torch_sync()
    dev.sync()
    t_start = time.perf_counter()
    for _ in range(steps):
        dev.step(dt, step, residual=False)
        step += 1
    dev.sync()
    torch_sync()
    elapsed = time.perf_counter() - t_start
    # per-kernel HIP-event timing (eager launches: events sit between the kernels)
    dev.kernel_times(reset=True)
    for _ in range(steps):
        dev.step(dt, step, residual=False)
        step += 1
    dev.sync()
    ms, _ = dev.kernel_times()
    log(f"[bench] timed {steps} steps ({'graph replay' if graphs else 'eager'}): "
        f"{elapsed:.3f}s; kernel ms slope={ms[0]:.3f} flux+relax={ms[1]:.3f}")
    st = dev.get_state(df=False)
    finite = bool(np.isfinite(st["w"]).all())
    cp = None
    if copy:
        cp = dev.copy_bandwidth(1 << 30, 10)
        log(f"[bench] plain copy of 1 GiB: {cp[1]:.3f} ms -> {cp[0]:.0f} GB/s")
    dev.close()
    return {"elapsed": elapsed, "ms": ms, "finite": finite, "copy": cp, "setup_s": t_setup,
            "state": st, "steps_run": step}


def form_errors(a, b, case):
    """Largest per-cell relative difference of rho, rho U, rho E (each cell's w scaled by its
    own max |w|) and of q (scaled by max |q| over the cells) between two runs of the same
    steps, over the fluid cells."""
    import numpy as np
    dim = case.config.solver.DIM
    nw = dim + 2
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    wa = np.asarray(a["w"]).reshape(-1, nw)[fluid]
    wb = np.asarray(b["w"]).reshape(-1, nw)[fluid]
    scale = np.max(np.abs(wb), axis=1)
    d = np.abs(wa - wb) / scale[:, None]
    qa = np.asarray(a["qf"]).reshape(-1, dim)[fluid]
    qb = np.asarray(b["qf"]).reshape(-1, dim)[fluid]
    return {"rho": float(d[:, 0].max()), "rhoU": float(d[:, 1:1 + dim].max()),
            "rhoE": float(d[:, nw - 1].max()),
            "q": float(np.abs(qa - qb).max() / np.abs(qb).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving stream and print its report")
    ap.add_argument("--no-other-form", action="store_true",
                    help="skip the run of the other storage form (FP64 beside FP32 and back)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--flux-waves", type=int, default=0,
                    help="waves per workgroup of the flux/relaxation kernel (0: library default)")
    ap.add_argument("--slope-vpt", type=int, default=0,
                    help="velocity cells per thread of the slope kernel (0: library default)")
    ap.add_argument("--workload", default="cylinder2d-vsamr", choices=list(WORKLOADS))
    ap.add_argument("--precision", default="fp32f", choices=["fp64", "fp32f"],
                    help="fp32f (f stored and advected in FP32, FP64 moments; the default) or "
                         "fp64 (the parity form)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="1: replay each step as one captured HIP graph (default); 0: eager")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: tune the launch options on the workload first (coordinate descent "
                         "over timed steps; the default), 0: library defaults")
    ap.add_argument("--vel-tiles", type=int, default=-1,
                    help="1/0: wave tiles of the velocity grid on/off (-1: library default)")
    ap.add_argument("--branches", type=int, default=0,
                    help="streams for the independent launches of a phase (0: library default)")
    ap.add_argument("--regular", type=int, default=-1,
                    help="1/0: specialised slope and gather of regular cells on/off "
                         "(-1: library default, on)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by physical-space domain decomposition, not run here)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    if args.serve_only:
        rep = serving_stream()
        log(f"[bench] serving: p50 {rep['p50_latency_s']:.3f}s p99 {rep['p99_latency_s']:.3f}s "
            f"{rep['updates_per_s'] / 1e9:.2f} G updates/s")
        print(json.dumps({"serving": rep}), flush=True)
        return

    t0 = time.perf_counter()
    case = make_case(args.workload, args.precision)
    adaptive = not int(case.topo.get("vs_uniform", 0))   # per-cell velocity grids
    DIM, NDF = case.config.solver.DIM, case.config.solver.NDF
    s_f = 8 if args.precision == "fp64" else 4
    B_SLOPE, B_FLUX_ITER, B_STEP = alg_bytes(s_f, adaptive, DIM, NDF)
    P = case.phase_cells
    sizes = np.diff(np.asarray(case.topo["vs_offset"]))[np.asarray(case.topo["bound_enc"]) >= 0]
    log(f"[bench] case built in {time.perf_counter() - t0:.1f}s: leaves={case.n_cells} "
        f"pairs={case.topo['n_pairs']} P={P} grid sizes {sizes.min()}-{sizes.max()}")
    tuned = None
    tune_rep = None
    if args.autotune:
        from kitamr_hip import autotune
        t1 = time.perf_counter()
        tuned, tms, trials = autotune.autotune(case, steps=3, warmup=1, log=log)
        tune_s = time.perf_counter() - t1
        log(f"[bench] autotune ({len(trials)} trials, {tune_s:.1f}s): {tuned} "
            f"{tms:.3f} ms/step")
        tune_rep = {"trials": [{"options": o, "ms_per_step_eager": m} for o, m in trials],
                    "seconds": tune_s,
                    "search": "coordinate descent over kitamr_hip.autotune.SPACE, 3 eager steps "
                              "per trial on a fresh context of the workload"}
    options = dict(tuned or {})
    if args.flux_waves:
        options["flux_waves"] = args.flux_waves
    if args.slope_vpt:
        options["slope_vpt"] = args.slope_vpt
    if args.vel_tiles >= 0:
        options["vel_tiles"] = args.vel_tiles
    if args.regular >= 0:
        options["regular"] = args.regular
    if args.branches:
        options["branches"] = args.branches
    r = timed_run(case, options, args.steps, args.warmup, bool(args.graphs), copy=True, log=log)
    copy_gbs, copy_ms = r["copy"]
    ms = r["ms"]
    elapsed = r["elapsed"]
    ms_step = 1e3 * elapsed / args.steps
    value = P * args.steps / elapsed
    kr_ms = ms[1]
    achieved = B_FLUX_ITER * P / (kr_ms * 1e-3) / 1e9
    slope_gbs = B_SLOPE * P / (ms[0] * 1e-3) / 1e9 if ms[0] > 0 else None
    step_gbs = B_STEP * value / 1e9
    out = {
        "metric": "phase-space updates/s + % HBM roofline (1 GPU)",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (f stored and advected; f64 moments)",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "precision": args.precision,
            "launch": "one HIP graph per step" if args.graphs else "eager launches",
            "launch_options": tuned or "library defaults",
            "leaves": case.n_cells,
            "phase_cells_per_step": P,
            "grid_cells_min_max": [int(sizes.min()), int(sizes.max())],
            "parallelism": "single GPU",
        },
        "hbm_roofline_pct": 100.0 * step_gbs / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flux_relax (fused flux + relaxation)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(args.workload, args.precision),
            "alg_bytes_per_update": B_FLUX_ITER,
            "avg_launch_ms": kr_ms,
        },
        "slope_kernel": {"avg_ms_per_step": ms[0], "alg_bytes_per_update": B_SLOPE,
                         "achieved_GBs": slope_gbs},
        "step_alg_bytes_per_update": B_STEP,
        "copy_bandwidth_GBs": copy_gbs,
        "state_finite": r["finite"],
        "setup_s": r["setup_s"],
    }
    if tune_rep:
        out["autotune"] = tune_rep
    if not args.no_other_form:
        other = "fp64" if args.precision == "fp32f" else "fp32f"
        co = with_precision(case, other)
        b_so, b_fio, b_sto = alg_bytes(8 if other == "fp64" else 4, adaptive, DIM, NDF)
        ro = timed_run(co, {}, args.steps, args.warmup, bool(args.graphs), log=log)
        vo = P * args.steps / ro["elapsed"]
        ao = b_fio * P / (ro["ms"][1] * 1e-3) / 1e9
        out[f"{other}_form"] = {
            "value": vo, "unit": "updates/s", "ms_per_step": 1e3 * ro["elapsed"] / args.steps,
            "hbm_roofline_pct": 100.0 * b_sto * vo / 1e9 / HBM_PEAK_GBS,
            "step_alg_bytes_per_update": b_sto,
            "k_flux_relax": {"avg_launch_ms": ro["ms"][1], "achieved_GBs": ao,
                             "frac": ao / HBM_PEAK_GBS, "alg_bytes_per_update": b_fio},
            "slope_kernel_ms": ro["ms"][0],
            "state_finite": ro["finite"],
        }
        out["speedup_fp32_vs_fp64"] = (value / vo) if other == "fp64" else (vo / value)
        if ro["steps_run"] == r["steps_run"]:
            e32, e64 = (r, ro) if other == "fp64" else (ro, r)
            out["fp32_form_vs_fp64_form"] = dict(
                form_errors(e32["state"], e64["state"], case),
                steps=r["steps_run"],
                note="per-cell relative difference after the same steps from the same state; "
                     "the FP64 form agrees with the FP64 oracle to 1e-12 (GPU parity tests)")
        log(f"[bench] {other} form: {vo / 1e9:.2f} G updates/s")
    if not args.no_serve:
        rep = serving_stream()
        log(f"[bench] serving: p50 {rep['p50_latency_s']:.3f}s p99 {rep['p99_latency_s']:.3f}s "
            f"{rep['updates_per_s'] / 1e9:.2f} G updates/s")
        out["serving"] = rep
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(case, max_steps=args.cpu_steps)
        cb = out["cpu_baseline"]
        out["speedup_vs_cpu"] = {"all_threads": value / cb["value"],
                                 "one_core": value / cb["one_core"]["value"]}
        log(f"[bench] cpu baseline {cb['value'] / 1e6:.1f} M updates/s "
            f"({cb['cores']} threads), {cb['one_core']['value'] / 1e6:.2f} M updates/s (1 thread)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
