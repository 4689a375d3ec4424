#!/usr/bin/env python3
"""Throughput of the SPH step on one MI355X: the 3D dam break at 1,024,000
fluid particles (80 x 80 x 160 at dx = 0.0075, 3-layer fixed boundary),
mixed precision (FP32 pair physics, FP64 positions/velocities) and FP64.

One "step" is one full leap-frog step of the hot path: reference-grid
bounding box, cell keys and sort, neighbour search, EOS, the fused force pass
and the integrator update, with every particle resident in HBM.

Prints ONE JSON line. ``value`` is the mixed-precision particle-steps/s
(fluid particles x steps / s); the FP64 figure, their ratio, the density
difference between them, the force kernel's roofline, the per-call rate of
the drop-in single_step with host buffers (PCIe and host work included), the
batched serving leg and a CPU baseline (the oracle restatement on the host
cores, on a bounded sample, calibrated against the compiled reference) ride
along; --general (or --only general) adds the general-path workloads
(elastic beam, elastic gate).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M 3D dam break, mixed vs FP64; p50/p99 batched-run latency"
PEAK_FP64 = 78.6     # TFLOP/s, MI355X vector FP64
PEAK_FP32 = 157.3    # TFLOP/s, MI355X FP32
FLOP_PER_PAIR_3D = 81  # algorithmic FLOP per interacting pair (3D dam-break physics set)
PEAK_HBM = 8.0e12    # B/s, MI355X HBM3E
# HBM bytes per force-kernel dispatch, from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
# scaled by the same counters on a plain 1 GiB copy (counters are not readable
# from inside a process, so the figure is recorded beside the kernel trace).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "force_kernel_hbm_traffic.json")
# Algorithmic bytes per particle-step of the whole step (each SoA field read
# and written once per kernel, neighbour re-reads excluded): neighbour search,
# fused pair kernel, integrator; and FLOP per particle-step outside the pairs.
STEP_BYTES = {"mixed": (190, 60, 250), "fp64": (217, 96, 328)}
FLOP_PER_PARTICLE = 40


def step_roofline(n_fluid, pairs, ms_per_step, precision):
    """Ideal step time T* = sum over the three stages of max(bytes / HBM peak,
    FLOP / VALU peak), against the measured step."""
    nnps, pair_b, integ = STEP_BYTES[precision]
    valu = (PEAK_FP32 if precision == "mixed" else PEAK_FP64) * 1e12
    t = [nnps * n_fluid / PEAK_HBM,
         max(pair_b * n_fluid / PEAK_HBM, FLOP_PER_PAIR_3D * pairs / valu),
         max(integ * n_fluid / PEAK_HBM, FLOP_PER_PARTICLE * n_fluid / valu)]
    t_star = 1e3 * sum(t)
    return {"t_star_ms": t_star, "ms_per_step": ms_per_step, "frac": t_star / ms_per_step,
            "stages_ms": {"nnps": 1e3 * t[0], "pair": 1e3 * t[1], "integrator": 1e3 * t[2]},
            "bytes_per_particle_step": nnps + pair_b + integ}


def measured_traffic(mixed):
    """HBM bytes per launch of the 3D fluid force kernel, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    tag = "k_force_fused<3, %s, false, false>" % ("true" if mixed else "false")
    for name, rec in table.items():
        if tag in name and rec.get("fetch_bytes") and rec.get("write_bytes") is not None:
            return rec["fetch_bytes"] + rec["write_bytes"]
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(case, precision, warmup, steps):
    from mysph_amd import config as C
    from mysph_amd.sphgpu import Simulation
    cfg = case["config"].copy()
    cfg.precision = precision
    sim = Simulation(case, cfg)
    sim.advance(warmup)   # returns after the context's stream has drained
    t0 = time.perf_counter()
    sim.advance(steps)
    t1 = time.perf_counter()
    adv_ms, force_ms, _ = sim.timing()
    stats = sim.neighbour_stats()
    state = {k: sim.get(k)[: case["n_real"]] for k in ("rho", "x", "v")}
    sim.close()
    return dict(wall_s=t1 - t0, event_ms=adv_ms, force_ms=force_ms, pairs=stats["pairs"],
                max_nbr=stats["max"], rho=state["rho"], state=state,
                label=C.FP64 if precision == 0 else C.MIXED)


def force_samples(case, precision, warmup, steps, every=10):
    """Average duration and pair count of the force pass over the timed
    steps. The timed run replays captured graphs, in which no launch can be
    timed alone, so a second run of the same workload advances in calls of
    `every` steps and times each call's closing force launch with HIP events
    on the context's stream; the pair count is that launch's."""
    from mysph_amd.sphgpu import Simulation
    cfg = case["config"].copy()
    cfg.precision = precision
    ms, pairs = [], []
    with Simulation(case, cfg) as sim:
        sim.advance(warmup)
        for _ in range(steps // every):
            sim.advance(every)
            ms.append(sim.timing()[1])
            pairs.append(sim.neighbour_stats()["pairs"])
    return float(np.mean(ms)), float(np.mean(pairs)), len(ms)


def pcie_inclusive(case, reps=3):
    """Per-step cost of the drop-in `single_step` at this size, where the
    host state crosses PCIe on every call: upload every particle array,
    evaluate the derivatives (search, EOS, force pass, fixed boundary), then
    download what the drop-in writes back for this fluid case, called without
    the solid-only Shear argument (dv/dt, de/dt, drho/dt, p, the two
    divergences). Host arrays are allocated once, as the drop-in's are; its
    Fortran packing loops are not included. FP64, the reference's arithmetic.
    The first call (context set-up) is untimed."""
    from mysph_amd import config as C
    from mysph_amd.sphgpu import Simulation, _dp
    names = ("dvdt", "dedt", "drhodt", "p", "divx", "divv")
    with Simulation(case) as sim:
        bufs = {k: np.zeros(sim.n * C.field_components(C.FIELD_NAMES[k], sim.dim)) for k in names}
        t = np.zeros((reps, 3))
        for r in range(reps + 1):
            t0 = time.perf_counter()
            sim.upload(case)      # each call returns after its stream has drained
            t1 = time.perf_counter()
            sim.derivatives()
            t2 = time.perf_counter()
            for k in names:
                sim._check(sim._L.sphgpu_download(sim._ctx, C.FIELD_NAMES[k], _dp(bufs[k])))
            t3 = time.perf_counter()
            if r:
                t[r - 1] = (t1 - t0, t2 - t1, t3 - t2)
        up = sum(a.nbytes for a in sim._keep.values() if a is not None)
    ms = 1e3 * t.mean(axis=0)
    step = float(ms.sum())
    return {"value": case["n_real"] / (step * 1e-3), "unit": "particle-steps/s",
            "ms_per_step": step, "upload_ms": float(ms[0]), "derivatives_ms": float(ms[1]),
            "download_ms": float(ms[2]), "bytes_up": int(up),
            "bytes_down": int(sum(b.nbytes for b in bufs.values())), "calls": reps,
            "path": "drop-in single_step per call: upload + derivatives + downloads, f64"}


def accuracy(case, nsteps, states=None, warm=10):
    """Mixed against FP64 on the GPU after nsteps from the same state (the FP64
    GPU path is checked against the oracle by the parity tests). `states`:
    the end states of the timed runs when they are nsteps long. A run made
    here is timed as well (steps warm+1 .. nsteps, after `warm` untimed
    steps), which gives the steady per-step figures of the whole workload
    whatever window the timed runs used."""
    from mysph_amd import config as C
    from mysph_amd.sphgpu import Simulation
    nf = case["n_real"]
    out = dict(states) if states else {}
    steady = {}
    for name, prec in (("fp64", C.FP64), ("mixed", C.MIXED)):
        if name in out:
            continue
        cfg = case["config"].copy()
        cfg.precision = prec
        with Simulation(case, cfg) as sim:
            sim.advance(warm)
            t0 = time.perf_counter()
            sim.advance(nsteps - warm)
            t1 = time.perf_counter()
            steady[name] = 1e3 * (t1 - t0) / (nsteps - warm)
            out[name] = {k: sim.get(k)[:nf] for k in ("rho", "x", "v")}
    rho_err = np.max(np.abs(out["mixed"]["rho"] - out["fp64"]["rho"]) / out["fp64"]["rho"])
    x_err = np.max(np.abs(out["mixed"]["x"] - out["fp64"]["x"])) / np.max(np.abs(out["fp64"]["x"]))
    v_err = np.max(np.abs(out["mixed"]["v"] - out["fp64"]["v"])) / np.max(np.abs(out["fp64"]["v"]))
    return {"steps": nsteps, "mixed_vs_fp64_rho_max_rel": float(rho_err),
            "mixed_vs_fp64_x_rel": float(x_err), "mixed_vs_fp64_v_rel": float(v_err)}, steady


def host_cpu():
    """The host CPU as lscpu reports it, and the cores this job may run on."""
    info = {}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = dict(line.split(":", 1) for line in txt.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        info["logical_cpus"] = int(kv.get("CPU(s)", "0") or 0)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        sockets = int(kv.get("Socket(s)", "0") or 0)
        info["physical_cores"] = cps * sockets or None
    except Exception as e:  # lscpu missing: the counts below still stand
        info["lscpu_error"] = repr(e)
    try:
        info["cpus_allowed"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["cpus_allowed"] = os.cpu_count() or 1
    return info


def baseline_threads(info):
    """Every core the job may use: the CPUs in its affinity mask, or fewer
    when the environment caps OpenMP (a shared GPU host allots a share of its
    cores to each GPU job)."""
    n = info.get("cpus_allowed") or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return n


def reference_calibration(threads, nsteps=100):
    """The compiled reference (its own modules, built from source into
    oracle/_ref/ref_driver) and the oracle port timed on the reference's own
    2D dam break (15,376 fluid + 5,734 wall particles), same steps, same
    threads: how the port's speed compares with the reference's."""
    import subprocess
    import tempfile
    import oracle
    from mysph_amd import cases
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    out = {"case": "reference 2D dam break (data/damBreak), 15376 fluid + 5734 walls",
           "steps": nsteps, "threads": threads}
    case = cases.reference_dam_break()
    oracle.set_threads(threads)
    sim = oracle.OracleSim(case)
    t0 = time.perf_counter()
    sim.run(nsteps)
    out["port_s"] = time.perf_counter() - t0
    out["port_value"] = case["n_real"] * nsteps / out["port_s"]
    if not os.path.exists(exe):
        out["reference_error"] = "oracle/_ref/ref_driver is not built"
        return out
    env = dict(os.environ, REF_THREADS=str(threads), OMP_NUM_THREADS=str(threads))
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        r = subprocess.run([exe, "dam_break", "run", str(nsteps), os.path.join(d, "o.bin")],
                           capture_output=True, text=True, env=env, cwd=d, timeout=600)
        dt = time.perf_counter() - t0
    if r.returncode != 0:
        out["reference_error"] = r.stderr[-500:]
        return out
    out["reference_s"] = dt
    out["reference_value"] = case["n_real"] * nsteps / dt
    out["port_over_reference"] = out["port_value"] / out["reference_value"]
    return out


def cpu_baseline(case, seconds_target=20.0):
    """The oracle (C restatement of the reference path, OpenMP) on the host,
    over a bounded number of full steps of the same workload, on every core
    the job may use; beside it, the compiled reference and the port timed on
    the reference's own case as a calibration."""
    import oracle
    info = host_cpu()
    threads = baseline_threads(info)
    oracle.set_threads(threads)
    sim = oracle.OracleSim(case)
    t0 = time.perf_counter()
    sim.run(1)                      # warm-up step (step 1)
    t1 = time.perf_counter()
    per = max(t1 - t0, 1e-3)
    nsteps = int(max(2, min(20, seconds_target / per)))
    t2 = time.perf_counter()
    sim.run(nsteps)
    t3 = time.perf_counter()
    del sim
    out = dict(value=case["n_real"] * nsteps / (t3 - t2), unit="particle-steps/s",
               cores=threads, kind="port",
               sample=f"the full workload ({case['n_real']} fluid + {case['n_dummy']} boundary), "
                      f"1 warm-up + {nsteps} timed leap-frog steps, oracle/sph_oracle.c with "
                      f"{threads} OpenMP threads",
               host=info)
    try:
        out["calibration"] = reference_calibration(threads)
    except Exception as e:  # the port's figure stands on its own
        out["calibration"] = {"error": repr(e)}
    return out


def general_leg(beam_steps=2000, gate_steps=100):
    """The general path's workloads (secondary lines, not the headline): the
    2D elastic beam (16,120 particles, artificial stress, clamped columns) and
    the 3D dam break with an elastic gate (2,046,000 fluid + 45,900 solid +
    467,742 fixed particles, physical viscosity, fluid-solid coupling, h per
    phase), each in FP64 and in mixed precision, timed after a warm-up."""
    from mysph_amd import cases
    from mysph_amd import config as C
    out = {}
    for name, make, warm, steps in (("elastic_beam_2d", cases.elastic_beam, 20, beam_steps),
                                    ("elastic_gate_3d", cases.dam_break_gate_3d, 10, gate_steps)):
        case = make()
        nt = case["n_real"]
        rec = {"real_particles": int(nt), "fixed_particles": int(case["n_dummy"]), "steps": steps,
               "warmup": warm}
        for pname, prec in (("fp64", C.FP64), ("mixed", C.MIXED)):
            r = run_gpu(case, prec, warm, steps)
            rec[pname] = {"ms_per_step": 1e3 * r["wall_s"] / steps, "particle_steps_per_s": nt * steps / r["wall_s"],
                          "force_pass_ms": r["force_ms"]}
        rec["mixed_over_fp64"] = rec["fp64"]["ms_per_step"] / rec["mixed"]["ms_per_step"]
        out[name] = rec
        log(f"{name}: fp64 {rec['fp64']['ms_per_step']:.3f} ms/step, mixed {rec['mixed']['ms_per_step']:.3f} "
            f"ms/step ({rec['mixed_over_fp64']:.2f}x)")
    return out


def serving_leg(args):
    """Continuous batching of the dam-break parameter sweep: tune the
    per-size tiling, find the saturation rate with all requests arriving at
    once, then serve Poisson streams at fractions of it."""
    import math
    from mysph_amd import config as C
    from mysph_amd import serving
    from mysph_amd.sphgpu import Batch

    t0 = time.perf_counter()
    reqs = serving.sweep_requests(args.serving_runs, steps=args.serving_steps)
    setup_s = time.perf_counter() - t0
    work = sum(r.n_fluid * r.steps for r in reqs)
    log(f"serving: {len(reqs)} requests, {sum(r.n_fluid for r in reqs)} fluid particles, "
        f"case set-up {setup_s:.1f} s")

    def make_batch(slots=None):
        return Batch(2, C.MIXED, slots=slots or args.serving_slots, cells=8 * (slots or args.serving_slots),
                     max_runs=256)

    tuner = serving.Tuner()
    reps = {0: 10000, 1: 40000, 2: 150000, 3: 450000}
    # candidates keep 3 cells per cutoff: the arena's cell room is sized for
    # that, and a finer grid (4 per cutoff: 1.8x the cells in 2D) that times
    # faster alone leaves room for fewer resident runs (17.4-17.7 runs/s at
    # saturation against 18.4 with 3 everywhere). Timed over steps 501-700,
    # where the flow has the speeds that set the rebuild rate of 2000-step runs.
    report = tuner.tune(make_batch, lambda k: serving.dam_break_request_case(reps[k], 1.25, 0.1),
                        candidates=((3, 0.15), (3, 0.08), (3, 0.25)),
                        fill=1_000_000, warmup=500, steps=200)
    log(f"serving: tuned tiling per size class {tuner.table}")

    def serve(rate):
        serving.poisson_arrivals(reqs, rate)
        with make_batch() as b:
            srv = serving.Server(b, chunk=args.serving_chunk, tuner=tuner)
            res = srv.serve(reqs)
        st = serving.latency_stats(res)
        st["advance_s"] = srv.advance_s
        st["chunks"] = srv.chunks
        return st

    burst = serve(math.inf)
    mean_work = work / len(reqs)
    sat_rate = burst["throughput"] / mean_work       # requests/s the GPU sustains
    log(f"serving: burst p50 {burst['p50_s']:.2f} s p99 {burst['p99_s']:.2f} s, "
        f"{burst['throughput']:.4g} particle-steps/s, saturation {sat_rate:.3g} runs/s")
    streams = []
    for f in args.serving_rates:
        st = serve(f * sat_rate)
        st.update(rate_factor=f, rate_runs_per_s=f * sat_rate)
        streams.append(st)
        log(f"serving: {f:g}x saturation: p50 {st['p50_s']:.2f} s p99 {st['p99_s']:.2f} s, "
            f"{st['throughput']:.4g} particle-steps/s")
    head = streams[0] if streams else burst
    return {
        "workload": f"{len(reqs)} 2D dam-break runs, fluid counts log-uniform in [5k, 500k] "
                    f"(seed {serving.SEED_SWEEP}), column aspect U[0.5,2], alpha U[0.05,0.2], "
                    f"{args.serving_steps} steps each, mixed precision, Poisson arrivals",
        "runs": len(reqs),
        "steps_per_run": args.serving_steps,
        "particle_steps_total": int(work),
        "arena_slots": args.serving_slots,
        "chunk_steps": args.serving_chunk,
        "tiling_per_size_class": {"bounds": list(tuner.bounds), "win_skin": tuner.table,
                                  "differs_by_class": len(set(map(tuple, tuner.table))) > 1,
                                  "margin_over_second": {str(k): (v[1][2] / v[0][2] - 1.0) if len(v) > 1 else None
                                                         for k, v in report.items()},
                                  "ns_per_particle_step": {str(k): v for k, v in report.items()}},
        "headline_rate_factor": head.get("rate_factor", "burst"),
        "p50_s": head["p50_s"],
        "p99_s": head["p99_s"],
        "throughput": head["throughput"],
        "burst": burst,
        "saturation_runs_per_s": sat_rate,
        "poisson": streams,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the workload is 1000 steps with steps 11-1000 timed
    ap.add_argument("--steps", type=int, default=990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.0075)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", choices=["both", "mixed", "fp64", "serving", "general"], default="both",
                    help="serving: only the batched serving leg; general: only the general-path workloads")
    ap.add_argument("--general", action="store_true",
                    help="also time the general-path workloads (elastic beam, elastic gate)")
    ap.add_argument("--general-steps", type=int, nargs=2, default=[2000, 100], metavar=("BEAM", "GATE"),
                    help="timed steps of the elastic beam and of the elastic gate")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="untimed run of both precisions to report the mixed density error")
    ap.add_argument("--no-serving", action="store_true", help="skip the batched serving leg")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the per-call upload/download rate of the drop-in single_step")
    ap.add_argument("--serving-runs", type=int, default=256)
    ap.add_argument("--serving-steps", type=int, default=2000)
    ap.add_argument("--serving-slots", type=int, default=8_000_000,
                    help="arena size in particle slots (admission waits for room)")
    ap.add_argument("--serving-chunk", type=int, default=50)
    ap.add_argument("--serving-rates", type=float, nargs="*", default=[0.5, 1.0],
                    help="Poisson arrival rates as fractions of the measured saturation rate")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not part of this build)",
              file=sys.stderr)
        sys.exit(2)

    if args.only == "serving":
        print(json.dumps({"serving": serving_leg(args)}), flush=True)
        return
    if args.only == "general":
        print(json.dumps({"general": general_leg(*args.general_steps)}), flush=True)
        return
    from mysph_amd import cases
    from mysph_amd import config as C
    case = cases.dam_break_3d(dx=args.dx)
    nf = case["n_real"]
    log(f"case: {nf} fluid + {case['n_dummy']} boundary particles")
    res = {}
    for name, prec in (("fp64", C.FP64), ("mixed", C.MIXED)):
        if args.only not in ("both", name):
            continue
        r = run_gpu(case, prec, args.warmup, args.steps)
        r["pss"] = nf * args.steps / r["wall_s"]
        r["force_ms"], r["pairs"], r["force_samples"] = force_samples(case, prec, args.warmup, args.steps)
        res[name] = r
        log(f"{name}: {r['pss']:.4g} particle-steps/s, {1e3 * r['wall_s'] / args.steps:.3f} ms/step, "
            f"force kernel {r['force_ms']:.3f} ms, pairs {r['pairs']}")
    head = res.get("mixed") or res["fp64"]
    peak = PEAK_FP32 if head is res.get("mixed") else PEAK_FP64
    flops = FLOP_PER_PAIR_3D * head["pairs"]
    achieved = flops / (head["force_ms"] * 1e-3) / 1e12 if head["force_ms"] > 0 else 0.0
    traffic = measured_traffic(head is res.get("mixed"))
    out = {
        "metric": METRIC,
        "value": head["pss"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair physics + f64 positions/velocities" if "mixed" in res else "f64",
        "data": "synthetic (seeded jittered lattice, no dataset)",
        "config": {
            "workload": f"3D dam break: 0.6x0.6x1.2 m column, 1.6x0.6 m tank with 1.3 m walls, "
                        f"dx={args.dx}, {nf} fluid + {case['n_dummy']} fixed boundary particles, "
                        "h=1.2dx cubic spline, Tait c0=10*sqrt(2gH), Monaghan alpha=0.1 beta=0, "
                        "continuity density, leap-frog, fixed dt",
            "fluid_particles": nf,
            "boundary_particles": case["n_dummy"],
            "dt": float(case["config"].dt),
        },
        "roofline": {
            "kernel": "k_force_fused (exact neighbour filter + continuity + pressure + Monaghan viscosity)",
            "bound": "valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "flop_per_pair": FLOP_PER_PAIR_3D,
            "pairs_per_launch": head["pairs"],
            "launch_ms": head["force_ms"],
            "launch_samples": head["force_samples"],
            "hbm_gbs_measured": (traffic / (head["force_ms"] * 1e-3) / 1e9
                                 if traffic and head["force_ms"] > 0 else None),
            "hbm_frac_measured": (traffic / (head["force_ms"] * 1e-3) / PEAK_HBM
                                  if traffic and head["force_ms"] > 0 else None),
        },
    }
    prec_head = "mixed" if "mixed" in res else "fp64"
    out["step_roofline"] = step_roofline(nf, head["pairs"], out["ms_per_step"], prec_head)
    if "fp64" in res:
        out["fp64"] = {"value": res["fp64"]["pss"],
                       "ms_per_step": 1e3 * res["fp64"]["wall_s"] / args.steps,
                       "force_kernel_ms": res["fp64"]["force_ms"]}
    if "fp64" in res and "mixed" in res:
        out["mixed_over_fp64"] = res["mixed"]["pss"] / res["fp64"]["pss"]
        d = np.max(np.abs(res["mixed"]["rho"] - res["fp64"]["rho"]) / res["fp64"]["rho"])
        out["mixed_rho_rel_err_vs_fp64"] = float(d)
    if "fp64" in res and "mixed" in res and args.accuracy_steps > 0:
        same = args.warmup + args.steps == args.accuracy_steps
        out["accuracy"], steady = accuracy(
            case, args.accuracy_steps,
            {k: res[k]["state"] for k in ("fp64", "mixed")} if same else None, warm=10)
        if same:
            steady = {k: 1e3 * res[k]["wall_s"] / args.steps for k in ("fp64", "mixed")}
        w = args.warmup if same else 10
        out["steady"] = {"steps": f"{w + 1}-{args.accuracy_steps}",
                         "mixed_ms": steady["mixed"], "fp64_ms": steady["fp64"],
                         "mixed_value": nf / (steady["mixed"] * 1e-3),
                         "fp64_value": nf / (steady["fp64"] * 1e-3),
                         "ratio": steady["fp64"] / steady["mixed"]}
        log(f"steady over steps {out['steady']['steps']}: mixed {steady['mixed']:.3f} ms, "
            f"fp64 {steady['fp64']:.3f} ms, ratio {out['steady']['ratio']:.3f}")
    if not args.no_pcie:
        out["pcie_inclusive"] = pcie_inclusive(case)
        log(f"drop-in single_step with host buffers: {out['pcie_inclusive']['ms_per_step']:.2f} ms/step")
    if args.general:
        out["general"] = general_leg(*args.general_steps)
    if not args.no_serving:
        out["serving"] = serving_leg(args)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(case)
        except Exception as e:  # the GPU figures stand on their own
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
