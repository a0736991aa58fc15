"""Benchmark of the MI355X forward-solver time loop.

Workload (one "step" of this benchmark = one full shot of the 2D elastic P-SV
configuration): 1000 x 1000 grid, dx 1 m, dt 1e-4 s, 4000 time steps, order-8
staggered operators, 30-cell CPML, 10 layers + 3 % Gaussian random
perturbations (seeds 2 / 102), one 30 Hz Ricker VY force source, 500 VX + 500
VY receivers.  Synthetic data, FP32 arithmetic, inputs resident in HBM before
the timed region.

Printed JSON line: metric = grid-cell updates/s of the tiled path (kernel
variant chosen by the autotuner); roofline of the dominant kernel (start/stop
events attached to each dispatch on the solver's stream);
the same workload on the reference-sequence GPU path ("plain", one kernel per
LAMA operation) for the batched-vs-plain ratio; and the CPU baseline: the
oracle (C restatement with the reference's operator sequence, OpenMP) on a
bounded sample of the same workload, timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "wave-simulation_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per 2D elastic cell-update: 5 fields read + written, 5 coefficients read (FP32)
BYTES_PER_CELL_UPDATE = 4 * (2 * 5 + 5)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed shots (each = 4000 time steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", type=int, default=50, help="time steps per hipGraph segment")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=3000,
                    help="time steps of the CPU baseline sample (about 10-15 s on 16 cores)")
    ap.add_argument("--tile-y", type=int, default=0, help="rows per tile (0 = autotune over all kernel variants)")
    ap.add_argument("--path", default="fused", choices=["fused", "phased", "strip"])
    ap.add_argument("--line", choices=["tmem", "adjoint", "3d", "batched-elastic"],
                    help="run only this extra line and print it as JSON (for profiling; no headline metric)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched TM-EM (64 shots), batched elastic (8 shots), adjoint (16 shots) and 3D elastic lines")
    return ap.parse_args()


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _kernel_stats(sol, L, wavesim, kinds=(0, 1, 2)):
    import ctypes as C

    out = []
    for kind in kinds:
        name = C.c_char_p()
        avg = C.c_double()
        n = C.c_longlong()
        b = C.c_double()
        wavesim._abi.check(L.wv_kernel_stats(sol._h, kind, C.byref(name), C.byref(avg), C.byref(n), C.byref(b)))
        out.append(dict(name=name.value.decode(), avg_ms=avg.value, launches=n.value, bytes=b.value))
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "el2d_strip_c2_1000x1000_hbm_bytes.md")
TRAFFIC_BATCHED = os.path.join(ROOT, "profiles", "el2d_strip_batched8_1000x1000_hbm_bytes.md")
TRAFFIC_3D = os.path.join(ROOT, "profiles", "el3d_c5_320cube_8shots_hbm_bytes.md")
TRAFFIC_TMEM = os.path.join(ROOT, "profiles", "tm2d_c3_800x400_64shots_hbm_bytes.md")


def _traffic(kernel, path=TRAFFIC_FILE):
    """HBM bytes per launch of `kernel` from a committed FETCH_SIZE/WRITE_SIZE summary (calibrated), or None."""
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6 and cells[0] == kernel:
                return float(cells[5].replace(",", "")) * 1e6
    except OSError:
        pass
    return None


def _roofline(kernels, measured_traffic=False, traffic_file=TRAFFIC_FILE):
    """Roofline of the dominant kernel; `traffic` from the committed counter summary of the same workload and
    kernel variant (null for a line or variant that has none)."""
    dom = max(kernels, key=lambda k: k["avg_ms"] * k["launches"])
    achieved = dom["bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(dom["name"], traffic_file) if measured_traffic else None, "kernel": dom["name"],
            "kernel_avg_us": round(dom["avg_ms"] * 1e3, 3), "bytes_per_launch": dom["bytes"]}


def bench_batched_tmem():
    """64 TM-EM shots of uneven receiver counts in one handle against the reference-sequence GPU path."""
    from wavesim import cases

    case = cases.tmem_c3()
    NT = case.settings["NT"]
    sol = cases.make_solver(case, steps_per_graph=50)
    choice = sol.tune(100)
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    t0 = time.perf_counter()
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    tb = time.perf_counter() - t0
    import wavesim

    L = wavesim.load()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, 300)
    sol.synchronize()
    kern = _kernel_stats(sol, L, wavesim)
    del sol
    plain = cases.make_solver(case, path="plain", shots=[case.shots[0]])
    ps = 300
    plain.run(0, 20)
    plain.synchronize()
    t1 = time.perf_counter()
    plain.run(20, 20 + ps)
    plain.synchronize()
    shot_plain = (time.perf_counter() - t1) * NT / ps
    n = len(case.shots)
    return {"workload": "2D TM-EM order 8, 800x400, 3000 steps, 64 shots (20-300 receivers each), one handle",
            "kernel": choice, "shots_per_s": round(n / tb, 3), "plain_shots_per_s": round(1.0 / shot_plain, 4),
            "ratio_vs_plain": round(n / tb * shot_plain, 1),
            "grid_cell_updates_per_s": round(case.cells * NT * n / tb, 1),
            "plain_sample": f"{ps} steps of one shot, scaled to {NT}",
            "roofline": _roofline([k for k in kern[:2] if k["launches"]],
                                  measured_traffic=choice == "phased-16", traffic_file=TRAFFIC_TMEM)}


def bench_adjoint():
    """Forward + adjoint gradient of 16 batched 2D elastic shots (600x600, 3000 steps), in the two memory
    plans: the FP16 record of every step (when it fits the device memory) and FP16 checkpoints every 50 steps
    with recomputation."""
    import numpy as np
    from wavesim import cases
    import wavesim

    case = cases.elastic_c4()
    NT = case.settings["NT"]
    n = len(case.shots)
    # the per-step record is written by the row-strip kernels: tune among their block heights
    sol = cases.make_solver(case, steps_per_graph=50, path="strip")
    choice = sol.tune(100)
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    sol.reset()
    sol.run(0, NT)
    t_fwd = sol.last_run_ms()
    L = wavesim.load()
    out = {"workload": "2D elastic order 8, 600x600, 3000 steps, 16 shots: forward + adjoint gradient "
                       "(FP64 gradient accumulation)",
           "forward_kernel": choice, "forward_ms": round(t_fwd, 2),
           "forward_shots_per_s": round(n / (t_fwd * 1e-3), 3)}
    for label, store in (("recorded", -1.0), ("checkpointed", 0.0)):
        sol.set_adjoint_store(store)
        sol.set_checkpointing(50, True)
        sol.run(0, NT)
        sol.synchronize()  # first run of the plan allocates and captures
        sol.reset()
        sol.run(0, NT)
        t_ck = sol.last_run_ms()
        for k in range(n):
            sol.set_adjoint_source(k, sol.seismograms(k))  # J = 1/2 ||d||^2
        sol.run_adjoint()
        t_adj = sol.last_run_ms()
        g = sol.gradient("vp")
        assert np.isfinite(g).all()
        # per-kernel split of the adjoint pass (dispatch-attached events, eager)
        wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
        sol.run_adjoint()
        sol.synchronize()
        kern = _kernel_stats(sol, L, wavesim, kinds=(3, 4, 5, 6))
        wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 0))
        plan = ("FP16 record of every step's derivative sums, no recomputation" if kern[0]["launches"] == 0
                else "FP16 checkpoints every 50 steps, segment recomputation")
        out[label] = {"plan": plan, "forward_with_plan_ms": round(t_ck, 2), "adjoint_ms": round(t_adj, 2),
                      "gradient_cost_vs_forward": round((t_ck + t_adj) / t_fwd, 3),
                      "gradient_shots_per_s": round(n / ((t_ck + t_adj) * 1e-3), 3),
                      "adjoint_kernels": [{"name": k["name"], "avg_us": round(k["avg_ms"] * 1e3, 2),
                                           "launches": k["launches"],
                                           "GB_per_s": round(k["bytes"] / max(k["avg_ms"], 1e-12) / 1e6, 1)}
                                          for k in kern if k["launches"]]}
    out["gradient_cost_vs_forward"] = out["recorded"]["gradient_cost_vs_forward"]
    return out


def bench_batched_elastic(wavesim, L, nshots=8):
    """The C2 model with 8 shots in one handle (sources 20 cells apart): the batched 2D elastic solver's
    grid-cell updates/s and the HBM fraction of its dominant kernel (out-of-cache working set)."""
    import copy

    from wavesim import cases

    case = cases.elastic_c2()
    NT = case.settings["NT"]
    sh0 = case.shots[0]
    case.shots = []
    for k in range(nshots):
        sh = copy.deepcopy(sh0)
        sh.src_index = sh0.src_index + 20 * k
        case.shots.append(sh)
    sol = cases.make_solver(case, steps_per_graph=50)
    choice = sol.tune(50)
    sol.run(0, 100)
    sol.synchronize()
    sol.reset()
    sol.run(0, NT)
    ms = sol.last_run_ms()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, 200)
    sol.synchronize()
    kern = _kernel_stats(sol, L, wavesim)
    return {"workload": f"2D elastic order 8, 1000x1000, 4000 steps, {nshots} shots batched in one handle",
            "kernel": choice, "grid_cell_updates_per_s": round(case.cells * NT * nshots / (ms * 1e-3), 1),
            "shots_per_s": round(nshots / (ms * 1e-3), 3), "ms_per_step": round(ms / NT, 4),
            "roofline": _roofline([k for k in kern[:2] if k["launches"]], measured_traffic=choice == "strip-16",
                                  traffic_file=TRAFFIC_BATCHED)}


def bench_3d(wavesim, L):
    """3D elastic, 320^3, 8 shots batched: 200 timed steps of the 2500-step configuration."""
    from wavesim import cases

    case = cases.elastic_c5()
    steps = 200
    sol = cases.make_solver(case, steps_per_graph=50)
    choice = sol.tune(20)
    sol.run(0, 50)
    sol.synchronize()
    sol.reset()
    sol.run(0, steps)
    ms = sol.last_run_ms()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, 20)
    sol.synchronize()
    kern = _kernel_stats(sol, L, wavesim)
    n = len(case.shots)
    return {"workload": "3D elastic order 8, 320^3, dx 2 m, 24-cell CPML, 8 shots batched (4320 receivers each); "
                        f"{steps} of 2500 steps timed",
            "kernel": choice, "grid_cell_updates_per_s": round(case.cells * steps * n / (ms * 1e-3), 1),
            "ms_per_step": round(ms / steps, 3),
            "roofline": _roofline(kern[:2], measured_traffic=choice == "3d-z4", traffic_file=TRAFFIC_3D)}


def cpu_baseline(case, nsteps):
    """Oracle (reference operator sequence, FP32, OpenMP) on a truncated run of the same workload."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import pyoracle as po
    from _cases import oracle_params

    threads = min(16, os.cpu_count() or 1)
    p = oracle_params(case, precision_threads=threads)
    sh = case.shots[0]
    args = (case.m0, case.m1, case.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type)
    t0 = time.perf_counter()
    po.run(p, *args, t_end=10)
    t1 = time.perf_counter()
    po.run(p, *args, t_end=10 + nsteps)
    t2 = time.perf_counter()
    loop = (t2 - t1) - (t1 - t0)  # setup and the first 10 steps cancel
    rate = case.cells * nsteps / loop
    return {"value": rate, "unit": "grid-cell updates/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} time steps of the same 1000x1000 order-8 CPML shot (loop only, setup subtracted)",
            "seconds": round(loop, 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(shots shard across GPUs with no in-loop communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import wavesim
    from wavesim import cases

    if args.line:
        L = wavesim.load()
        one = {"tmem": bench_batched_tmem, "adjoint": bench_adjoint, "3d": lambda: bench_3d(wavesim, L),
               "batched-elastic": lambda: bench_batched_elastic(wavesim, L)}
        print(json.dumps({args.line: one[args.line]()}))
        return

    case = cases.elastic_c2()
    NT = case.settings["NT"]
    sol = cases.make_solver(case, steps_per_graph=args.graph, tile_y=args.tile_y, path=args.path)
    choice = sol.tune(200) if (args.tile_y == 0 and args.path == "fused") else f"{args.path}-{args.tile_y}"

    for _ in range(args.warmup):
        sol.reset()
        sol.run(0, NT)
    sol.synchronize()
    t0 = time.perf_counter()
    gpu_ms = 0.0
    for _ in range(args.steps):
        sol.reset()
        sol.run(0, NT)
        gpu_ms += sol.last_run_ms()
    sol.synchronize()
    wall = time.perf_counter() - t0
    seis = sol.seismograms(0)
    assert np.isfinite(seis).all() and np.abs(seis).max() > 0

    updates = case.cells * NT * args.steps
    value = updates / wall
    ms_per_step = wall * 1e3 / args.steps
    step_bytes = BYTES_PER_CELL_UPDATE * case.cells * NT

    # per-kernel timing pass: eager launches with start/stop events attached to each dispatch
    L = wavesim.load()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    kernels = _kernel_stats(sol, L, wavesim)
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 0))
    roofline = _roofline(kernels, measured_traffic=choice == "strip-16")

    out = {
        "metric": "grid-cell updates/s (2D elastic order-8 fused, 1000x1000, 4000 steps)",
        "value": round(value, 1),
        "unit": "grid-cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "2D elastic P-SV order 8, 1000x1000, dx 1 m, dt 1e-4 s, 4000 steps, 30-cell CPML, "
                               "10 layers + 3% Gaussian perturbation, 1 VY Ricker 30 Hz shot, 1000 receivers",
                   "path": args.path, "kernel": choice, "steps_per_graph": args.graph, "shots_per_step": 1},
        "hbm_frac_step": round(step_bytes * args.steps / wall / 1e9 / HBM_PEAK_GBS, 4),
        "gpu_event_ms_per_step": round(gpu_ms / args.steps, 3),
        "kernels": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in d.items()} for d in kernels],
        "roofline": roofline,
    }

    if not args.no_plain:
        ps = args.plain_steps
        plain = cases.make_solver(case, path="plain")
        plain.run(0, 20)
        plain.synchronize()
        t1 = time.perf_counter()
        plain.run(20, 20 + ps)
        plain.synchronize()
        tp = time.perf_counter() - t1
        plain_shot_s = tp * NT / ps
        out["plain"] = {"shots_per_s": round(1.0 / plain_shot_s, 4), "fused_shots_per_s": round(args.steps / wall, 4),
                        "ratio": round((args.steps / wall) * plain_shot_s, 2),
                        "sample": f"{ps} time steps, scaled to {NT}"}

        del plain
    del sol
    _progress(f"C2: {value:.4g} cell-updates/s, {ms_per_step:.1f} ms per shot")
    if not args.no_extras:
        out["batched_tmem_64_shots"] = bench_batched_tmem()
        _progress(f"batched TM-EM: {out['batched_tmem_64_shots']['shots_per_s']} shots/s")
        out["batched_elastic_8_shots"] = bench_batched_elastic(wavesim, L)
        _progress(f"batched elastic: {out['batched_elastic_8_shots']['grid_cell_updates_per_s']:.4g} cell-updates/s")
        out["adjoint_16_shots"] = bench_adjoint()
        _progress(f"adjoint: cost {out['adjoint_16_shots']['recorded']['gradient_cost_vs_forward']} x forward "
                  f"(recorded), {out['adjoint_16_shots']['checkpointed']['gradient_cost_vs_forward']} x (checkpointed)")
        out["elastic3d_8_shots"] = bench_3d(wavesim, L)
        _progress(f"3D: {out['elastic3d_8_shots']['grid_cell_updates_per_s']:.4g} cell-updates/s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
        _progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
