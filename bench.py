"""Benchmark of the MI355X forward-solver time loop.

Workload (one "step" of this benchmark = one full shot of the 2D elastic P-SV
configuration): 1000 x 1000 grid, dx 1 m, dt 1e-4 s, 4000 time steps, order-8
staggered operators, 30-cell CPML, 10 layers + 3 % Gaussian random
perturbations (seeds 2 / 102), one 30 Hz Ricker VY force source, 500 VX + 500
VY receivers.  Synthetic data, FP32 arithmetic, inputs resident in HBM before
the timed region.

Printed JSON line: value = grid-cell updates/s of the hand-written kernels
(kernel variant chosen by the autotuner); `roofline` = algorithmic HBM bytes of
one time step (60 B per cell-update for 2D elastic plus the CPML memory
variables of the frame cells) over the GPU time of one time step's kernels
(start/stop events attached to each dispatch on the solver's stream); the same
workload on the reference-sequence GPU path ("plain", one kernel per LAMA
operation) for the shots/s ratio; `parity`: the GPU seismograms against the
oracle's float and double runs of the whole shot; and the CPU baseline: the
oracle (C restatement with the reference's operator sequence, OpenMP) and its
fused-loop mode on a bounded sample of the same workload, timed on this host's
cores (threads pinned, median of 3).
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
METRIC = "grid-cell updates/s (% HBM peak); batched+adjoint shots/s vs plain at 1e-5"

# Algorithmic HBM bytes of one time step (FP32; every field read and written once, every
# coefficient read once and shared by the B shots of a batch; CPML memory variables read
# and written on the frame cells only, 8 B each per axis):
#   2D acoustic 4 (2*3 + 3/B) per cell + 2 memory variables per frame cell and axis
#   2D elastic  4 (2*5 + 5/B)            + 4
#   2D TM-EM    4 (2*3 + 4/B)            + 2
#   3D elastic  4 (2*9 + 8/B)            + 6
PHYS_BYTES = {"acoustic": (3, 3, 2), "elastic2d": (5, 5, 4), "tmem": (3, 4, 2), "elastic3d": (9, 8, 6)}


def alg_bytes_per_step(kind, NX, NY, NZ=1, B=1, W=0):
    nf, nc, npsi = PHYS_BYTES[kind]
    cells = NX * NY * NZ
    if NZ > 1:
        frame = 2 * W * (NY * NZ + NX * NZ + NX * NY)
    else:
        frame = 2 * W * (NY + NX)
    return B * cells * 4 * (2 * nf + nc / B) + B * npsi * 8 * frame


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed shots (each = 4000 time steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", type=int, default=50, help="time steps per hipGraph segment")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=1000,
                    help="time steps of each CPU baseline sample (a few seconds on 16 cores)")
    ap.add_argument("--tile-y", type=int, default=0, help="rows per tile (0 = autotune over all kernel variants)")
    ap.add_argument("--path", default="fused", choices=["fused", "phased", "strip"])
    ap.add_argument("--line", choices=["tmem", "adjoint", "3d", "batched-elastic"],
                    help="run only this extra line and print it as JSON (for profiling; no headline metric)")
    ap.add_argument("--variant", default="",
                    help="with --line: kernel variant instead of the autotuner's pick (e.g. phased-16, strip-16, "
                         "3d-z4), so a profiled run is reproducible")
    ap.add_argument("--cpu-samples", type=int, default=3, help="CPU baseline samples (median reported)")
    ap.add_argument("--no-parity", action="store_true", help="skip the whole-shot oracle parity runs")
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


def _step_roofline(kernels, steps, alg_bytes, traffic_file=None):
    """Roofline of one time step: the algorithmic bytes of a step over the GPU time of the step's kernels
    (dispatch-attached events, summed over the kernels of `steps` eager time steps); `traffic` = the HBM bytes
    the step's kernels move, from the committed counter summary of the same workload and kernel variant (null
    when there is none)."""
    run = [k for k in kernels if k["launches"]]
    step_ms = sum(k["avg_ms"] * k["launches"] for k in run) / steps
    achieved = alg_bytes / (step_ms * 1e-3) / 1e9
    main = [k for k in run if k["launches"] >= steps]  # the per-step kernels (not the end-of-run flush)
    traffic = None
    if traffic_file:
        tr = [_traffic(k["name"], traffic_file) for k in main]
        traffic = sum(tr) if tr and all(t is not None for t in tr) else None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": " + ".join(k["name"] for k in main) + " (one time step)",
            "step_us": round(step_ms * 1e3, 3), "alg_bytes_per_step": alg_bytes,
            "kernel_model": [{"name": k["name"], "avg_us": round(k["avg_ms"] * 1e3, 3), "launches": k["launches"],
                              "model_bytes_per_launch": k["bytes"]} for k in run]}


VARIANTS = {"phased-8": ("phased", 8), "phased-16": ("phased", 16), "phased-32": ("phased", 32),
            "fused-8": ("fused", 8), "fused-16": ("fused", 16), "fused-24": ("fused", 24),
            "strip-8": ("strip", 8), "strip-16": ("strip", 16), "strip-32": ("strip", 32),
            "persistent-8": ("persistent", 8), "persistent-16": ("persistent", 16), "persistent-32": ("persistent", 32),
            "3d-z4": ("fused", 4), "3d-z8": ("fused", 8), "3d-z16": ("fused", 16)}


def _solver(case, variant, tune_steps, **kw):
    """A prepared solver on `variant` (a key of VARIANTS) or on the autotuner's pick ('')."""
    from wavesim import cases

    if variant:
        path, ty = VARIANTS[variant]
        kw = dict(kw, path=kw.get("path", path) if path == "fused" else path, tile_y=ty)
        return cases.make_solver(case, **kw), variant
    sol = cases.make_solver(case, **kw)
    return sol, sol.tune(tune_steps)


def bench_batched_tmem(variant="", plain_leg=True):
    """64 TM-EM shots of uneven receiver counts in one handle against the reference-sequence GPU path."""
    from wavesim import cases
    import wavesim

    case = cases.tmem_c3()
    NT = case.settings["NT"]
    sol, choice = _solver(case, variant, 100, steps_per_graph=50)
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    t0 = time.perf_counter()
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    tb = time.perf_counter() - t0
    L = wavesim.load()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, 300)
    sol.synchronize()
    kern = _kernel_stats(sol, L, wavesim)
    del sol
    n = len(case.shots)
    s = case.settings
    out = {"workload": "2D TM-EM order 8, 800x400, 3000 steps, 64 shots (20-300 receivers each), one handle",
           "kernel": choice, "shots_per_s": round(n / tb, 3),
           "grid_cell_updates_per_s": round(case.cells * NT * n / tb, 1),
           "roofline": _step_roofline(kern, 300, alg_bytes_per_step("tmem", s["NX"], s["NY"], 1, n,
                                                                    s["boundary_width"]),
                                      TRAFFIC_TMEM if choice == "phased-16" else None)}
    if plain_leg:
        plain = cases.make_solver(case, path="plain", shots=[case.shots[0]])
        ps = 300
        plain.run(0, 20)
        plain.synchronize()
        t1 = time.perf_counter()
        plain.run(20, 20 + ps)
        plain.synchronize()
        shot_plain = (time.perf_counter() - t1) * NT / ps
        out.update({"plain_shots_per_s": round(1.0 / shot_plain, 4), "ratio_vs_plain": round(n / tb * shot_plain, 1),
                    "plain_sample": f"{ps} steps of one shot, scaled to {NT}"})
    return out


def bench_adjoint(variant=""):
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
    sol, choice = _solver(case, variant, 100, steps_per_graph=50, path="strip")
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


def bench_batched_elastic(wavesim, L, nshots=8, variant=""):
    """The C2 model with 8 shots in one handle (sources 20 cells apart): the batched 2D elastic solver's
    grid-cell updates/s and its step roofline (out-of-cache working set)."""
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
    sol, choice = _solver(case, variant, 50, steps_per_graph=50)
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
    s = case.settings
    return {"workload": f"2D elastic order 8, 1000x1000, 4000 steps, {nshots} shots batched in one handle",
            "kernel": choice, "grid_cell_updates_per_s": round(case.cells * NT * nshots / (ms * 1e-3), 1),
            "shots_per_s": round(nshots / (ms * 1e-3), 3), "ms_per_step": round(ms / NT, 4),
            "roofline": _step_roofline(kern, 200, alg_bytes_per_step("elastic2d", s["NX"], s["NY"], 1, nshots,
                                                                     s["boundary_width"]),
                                       TRAFFIC_BATCHED if choice == "strip-16" else None)}


def bench_3d(wavesim, L, variant=""):
    """3D elastic, 320^3, 8 shots batched: 200 timed steps of the 2500-step configuration."""
    from wavesim import cases

    case = cases.elastic_c5()
    steps = 200
    sol, choice = _solver(case, variant, 20, steps_per_graph=50)
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
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 0))
    n = len(case.shots)
    s = case.settings
    out = {"workload": "3D elastic order 8, 320^3, dx 2 m, 24-cell CPML, 8 shots batched (4320 receivers each); "
                       f"{steps} of 2500 steps timed",
           "kernel": choice, "grid_cell_updates_per_s": round(case.cells * steps * n / (ms * 1e-3), 1),
           "ms_per_step": round(ms / steps, 3),
           "roofline": _step_roofline(kern, 20, alg_bytes_per_step("elastic3d", s["NX"], s["NY"], s["NZ"], n,
                                                                   s["boundary_width"]),
                                      TRAFFIC_3D if choice == "3d-z4" else None)}
    del sol
    out["adjoint"] = bench_3d_adjoint(wavesim, choice, steps)
    return out


def bench_3d_adjoint(wavesim, variant, steps, nshots=4, every=50):
    """Forward + adjoint gradient of the 320^3 case over its first `steps` steps (misfit 1/2 ||d||^2
    of every shot), FP16 checkpoints every `every` steps, the memory plan the budget gives (3/4 of
    the free device memory).  Four shots per handle: at eight, the FP32 record of one step
    (9 fields, 10.9 GB) and the sub-segment states (13.7 GB each) next to the adjoint's own
    fields leave no plan within 288 GB at a 50-step interval."""
    from wavesim import cases
    import numpy as np

    case = cases.elastic_c5(nshots=nshots, NT=steps)
    sol, _ = _solver(case, variant, 0, steps_per_graph=50)
    sol.run(0, 50)
    sol.synchronize()
    sol.reset()
    sol.run(0, steps)
    ms = sol.last_run_ms()
    sol.set_checkpointing(every, True)
    k2, nsub, host = sol.adjoint_plan()
    sol.run(0, steps)
    sol.synchronize()
    t_ck = sol.last_run_ms()
    for k in range(nshots):
        sol.set_adjoint_source(k, sol.seismograms(k))
    sol.run_adjoint()
    t_adj = sol.last_run_ms()
    assert np.isfinite(sol.gradient("vp")).all()
    return {"workload": f"the same grid and model, {nshots} of its shots in one handle, the first {steps} steps",
            "plan": {"checkpoint_every": every, "fp16": True, "sub_segment_steps": k2, "sub_segments": nsub,
                     "checkpoints_in_host_memory": host},
            "forward_ms": round(ms, 2), "forward_with_checkpoints_ms": round(t_ck, 2),
            "adjoint_ms": round(t_adj, 2), "gradient_cost_vs_forward": round((t_ck + t_adj) / ms, 3)}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_leg(case, nsteps, nsamples, gpu_seis, parity=True):
    """CPU baseline and the whole-shot parity check, both on the oracle (the only leg that runs it).

    Baseline: the oracle's operator-sequence mode (the reference's run() sequence: CSR SpMV per
    derivative, separate vector passes, CPML vectors; FP32, OpenMP) and its fused-loop mode (same
    per-cell arithmetic, two passes per step), each timed on `nsteps` steps of the same shot with its
    setup cancelled (a 10-step run subtracted), median of `nsamples`, threads pinned to cores.
    Parity: the GPU seismograms against the oracle's FP32 and FP64 runs of the whole shot."""
    import statistics

    import numpy as np

    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import pyoracle as po
    from _cases import oracle_params, oracle_threads

    threads = oracle_threads()
    p = oracle_params(case, precision_threads=threads)
    sh = case.shots[0]
    args = (case.m0, case.m1, case.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type)

    def loop_seconds(fn):
        t0 = time.perf_counter()
        fn(10)
        t1 = time.perf_counter()
        fn(10 + nsteps)
        t2 = time.perf_counter()
        return (t2 - t1) - (t1 - t0)

    seq = statistics.median(loop_seconds(lambda te: po.run(p, *args, t_end=te)) for _ in range(nsamples))
    fused = statistics.median(loop_seconds(lambda te: po.run_fused_elastic2d(p, *args, t_end=te))
                              for _ in range(nsamples))
    base = {"value": round(case.cells * nsteps / seq, 1), "unit": "grid-cell updates/s", "cores": threads,
            "kind": "port", "cpu": _cpu_model(), "threads_pinned": os.environ.get("OMP_PROC_BIND"),
            "sample": f"{nsteps} time steps of the same 1000x1000 order-8 CPML shot, operator-sequence mode "
                      f"(loop only, setup subtracted), median of {nsamples}",
            "seconds": round(seq, 3),
            "fused_loop": {"value": round(case.cells * nsteps / fused, 1), "unit": "grid-cell updates/s",
                           "cores": threads, "seconds": round(fused, 3),
                           "sample": f"{nsteps} time steps, fused-loop mode, median of {nsamples}"}}
    if not parity:
        return base, None

    def rel(a, b):
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))

    t0 = time.perf_counter()
    o32, _ = po.run(p, *args, precision="f32")
    # FP64: the fused-loop mode (its results equal the operator-sequence mode's to the last bit)
    o64 = po.run_fused_elastic2d(p, *args, precision="f64")
    par = {"relL2_vs_f32_oracle": rel(gpu_seis, o32), "relL2_vs_f64_oracle": rel(gpu_seis, o64),
           "oracle_f32_vs_f64": rel(o32, o64), "tolerance": 1e-5,
           "sample": "the whole shot (4000 steps, 1000 traces)", "seconds": round(time.perf_counter() - t0, 1)}
    return base, par


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots would shard across GPUs, one handle per GPU, with no communication inside the "
              "time loop)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import wavesim
    from wavesim import cases

    if args.line:
        L = wavesim.load()
        v = args.variant
        one = {"tmem": lambda: bench_batched_tmem(v, plain_leg=not args.no_plain),
               "adjoint": lambda: bench_adjoint(v), "3d": lambda: bench_3d(wavesim, L, v),
               "batched-elastic": lambda: bench_batched_elastic(wavesim, L, variant=v)}
        print(json.dumps({args.line: one[args.line]()}))
        return

    case = cases.elastic_c2()
    NT = case.settings["NT"]
    s = case.settings
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
    sol.check_finite()  # Simulation.cpp:519
    assert np.abs(seis).max() > 0

    updates = case.cells * NT * args.steps
    value = updates / wall
    ms_per_step = wall * 1e3 / args.steps
    step_bytes = alg_bytes_per_step("elastic2d", s["NX"], s["NY"], 1, 1, s["boundary_width"])

    # per-kernel timing pass: eager launches with start/stop events attached to each dispatch
    L = wavesim.load()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.reset()
    sol.run(0, NT)
    sol.synchronize()
    kernels = _kernel_stats(sol, L, wavesim)
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 0))
    roofline = _step_roofline(kernels, NT, step_bytes, TRAFFIC_FILE if choice == "strip-16" else None)

    out = {
        "metric": METRIC,
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
                               "10 layers + 3% Gaussian perturbation, 1 VY Ricker 30 Hz shot, 1000 receivers "
                               "(one benchmark step = one shot)",
                   "path": ("hand-written kernels, autotuned over single-pass, two-launch and row-strip variants"
                            if args.path == "fused" and args.tile_y == 0 else args.path),
                   "kernel": choice,
                   "launches_per_time_step": 1 if choice.startswith("fused") else 2,
                   "steps_per_graph": args.graph, "shots_per_step": 1},
        # the step's algorithmic bytes over the whole timed loop (graphs, flushes, resets included)
        "hbm_frac_step": round(step_bytes * NT * args.steps / wall / 1e9 / HBM_PEAK_GBS, 4),
        "gpu_event_ms_per_step": round(gpu_ms / args.steps, 3),
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
        out["plain"] = {"shots_per_s": round(1.0 / plain_shot_s, 4), "hand_written_shots_per_s": round(args.steps / wall, 4),
                        "ratio": round((args.steps / wall) * plain_shot_s, 2),
                        "sample": f"{ps} time steps, scaled to {NT}"}

        del plain
    del sol
    _progress(f"C2: {value:.4g} cell-updates/s, {ms_per_step:.1f} ms per shot, step roofline {roofline['frac']}")
    if not args.no_extras:
        out["batched_tmem_64_shots"] = bench_batched_tmem()
        _progress(f"batched TM-EM: {out['batched_tmem_64_shots']['shots_per_s']} shots/s")
        out["batched_elastic_8_shots"] = bench_batched_elastic(wavesim, L)
        _progress(f"batched elastic: {out['batched_elastic_8_shots']['grid_cell_updates_per_s']:.4g} cell-updates/s")
        out["adjoint_16_shots"] = bench_adjoint()
        _progress(f"adjoint: cost {out['adjoint_16_shots']['recorded']['gradient_cost_vs_forward']} x forward "
                  f"(recorded), {out['adjoint_16_shots']['checkpointed']['gradient_cost_vs_forward']} x (checkpointed)")
        out["elastic3d_8_shots"] = bench_3d(wavesim, L)
        _progress(f"3D: {out['elastic3d_8_shots']['grid_cell_updates_per_s']:.4g} cell-updates/s, gradient cost "
                  f"{out['elastic3d_8_shots']['adjoint']['gradient_cost_vs_forward']} x forward")
    if not args.no_cpu:
        base, par = cpu_baseline_leg(case, args.cpu_steps, args.cpu_samples, seis, parity=not args.no_parity)
        out["cpu_baseline"] = base
        if par is not None:
            out["parity"] = par
            _progress(f"parity: relL2 vs f32 oracle {par['relL2_vs_f32_oracle']:.2e}, "
                      f"vs f64 {par['relL2_vs_f64_oracle']:.2e}")
        _progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
