#!/usr/bin/env python3
"""Walker-steps/s of the MI355X walker engine on the L=32, D=32, 2048-walker
Heisenberg chain (FP32 contractions; QR, sampling, weights, overlaps and the
energy estimator in FP64), next to the reference algorithm on the host CPU.

One step = one full Walk step of MPSQMC2 (MPSQMC2.cpp:465-511): propagation of
every walker through all 31 couplings with normalisation, the local energy of
every walker, the weight/energy reductions and population control.  Inputs are
resident in HBM before the timed region; the device RNG supplies the
auxiliary-field draws.

The second figure of the metric, variance x GPU-s, is the blocking-analysis
variance of the mean projected energy over a fixed number of further steps
(--var-steps, whatever --steps is) times their wall time (the cost of a run at
a fixed statistical error is proportional to it), for the reference proposal
and for the learned one.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per walker-step of the environment-sweep algorithm at
# L=32, D=32 (d=2, 31 couplings): FP32 contraction flops and FP64 QR flops.
F32_PER_WS = 61.5e6
F64_PER_WS = 15.9e6
# The reference algorithm's work per walker-step on the same workload (16 + 1
# full-chain transfer-matrix overlaps and a full QR left-normalisation per
# coupling, MPSQMC2.cpp:540-600, plus the energy through H|trial>): the FP64
# flops the CPU baseline does per walker-step.
F_REF_PER_WS = 3.506e9
PEAK_FP32 = 157.3e12   # MI355X dense FP32 matrix
PEAK_FP64 = 78.6e12    # MI355X FP64

# Open-boundary snake lattices on the general path (long-range couplings):
# trial bond D_T, walker bonds drawn uniformly from D_W (the reference runs one
# D_W per job), and the environment-sweep FP32 / FP64 flops per walker-step
# for each walker bond.
SNAKES = {
    "snake4x4": dict(rows=4, cols=4, DT=64, DW=(16, 32, 48, 64), N=4096,
                     F32={16: 84.5e6, 32: 200e6, 48: 325e6, 64: 474e6},
                     F64={16: 1.8e6, 32: 11e6, 48: 29e6, 64: 62e6}),
    "snake6x6": dict(rows=6, cols=6, DT=128, DW=(128,), N=1024, F32={128: 25.7e9}, F64={128: 3.0e9}),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="chain", choices=["chain"] + sorted(SNAKES),
                    help="chain: the L=32, D=32 chain (default); snake4x4 / snake6x6: the 2D snake lattices "
                         "on the general path")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (chain 2048, snakes 64)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (chain 200, snakes 1)")
    ap.add_argument("--walkers", type=int, default=2048)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--D", type=int, default=32)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-walkers", type=int, default=160,
                    help="walkers of the CPU baseline's timed steps (a bounded sample of the workload)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed CPU steps after one warm-up step")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: one per physical core (at most 16)")
    ap.add_argument("--var-steps", type=int, default=1024,
                    help="steps of the variance x GPU-s runs (plain and learned), independent of --steps")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay timing")
    ap.add_argument("--no-autotune", action="store_true", help="snakes: keep the default tile table")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=20131015)
    ap.add_argument("--random-trial", action="store_true", help="random trial instead of the DMRG one")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal comparison")
    ap.add_argument("--learned-samples", type=int, default=1_000_000)
    ap.add_argument("--learned-epochs", type=int, default=16)
    ap.add_argument("--learned-mix", type=float, default=0.01,
                    help="weight of the uniform part of the learned proposal (keeps its support)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB ten times (reference for HBM counters)")
    args = ap.parse_args()
    chain = args.workload == "chain"
    if args.steps is None:
        args.steps = 2048 if chain else 64
    if args.warmup is None:
        args.warmup = 200 if chain else 1
    return args


def calibrate_copy(mib):
    """A plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of its dispatches
    against mib MiB read and written give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(10):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"bench: copied {mib} MiB ten times")


TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "walker32_kernel_chain_L32_D32_N2048_hbm.md")


def measured_traffic(L, D, N, walkers_per_launch):
    """HBM bytes per walker-kernel launch: the bytes per walker-step measured
    with rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by the copy-kernel
    reference) in the committed counter summary, times the walkers of an
    average launch of this run; None for other workloads."""
    if (L, D, N) != (32, 32, 2048) or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        if line.startswith("traffic_bytes_per_walker_step:"):
            return float(line.split(":", 1)[1]) * walkers_per_launch
    return None


def make_trial(system, D, seed, random_trial=False):
    """The trial state: two-site DMRG of the chain in the D bond profile (the
    reference's MPSQMC2 DT constructor does the same), or a random
    left-normalised MPS of that profile."""
    from mpsqmc_amd import mps as M
    if random_trial:
        vd, data = M.random_mps(system.L, D, 2, seed)
        data, _ = M.left_normalize(vd, data)
        return vd, data, None
    vd, data, E = M.dmrg(system, D, max_sweeps=6, tol=1e-8, seed=seed)
    return vd, data, E


def physical_cores():
    """Physical cores available to this process (lscpu's core/socket counts,
    bounded by the CPU affinity mask)."""
    import subprocess
    n_aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = {ln.strip() for ln in out.splitlines() if ln and not ln.startswith("#")}
        return max(1, min(len(cores), n_aff)), len(cores), n_aff
    except Exception:
        return n_aff, None, n_aff


def cpu_baseline(system, vd, data, walkers, steps, threads):
    """The reference algorithm (16 + 1 full-chain overlaps per coupling, a full
    QR left-normalisation after every operator, energy through H|trial>) in the
    FP64 C restatement, OpenMP over walkers: one untimed warm-up step, then
    `steps` timed Walk steps (propagation, energy, population control) on a
    bounded sample of the walkers."""
    import oracle
    use, phys, aff = physical_cores()
    threads = min(threads or use, 16, aff)
    pop = oracle.Population(system.L, 2, vd, data, system.J, system.h, system.dtau, system.tables_distinct(),
                            system.grid().reshape(-1), 4, build_terms=True, threads=threads, energy_mode=2)
    rng = np.random.default_rng(1)
    pop.init_walkers(min(walkers, 4 * threads), int(vd.max()))
    pop.propagate(rng.uniform(size=pop.nwalkers * pop.nC))  # warm-up step (page-in, thread start)
    pop.energy(1, True)
    pop.init_walkers(walkers, int(vd.max()))
    ws = 0
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        n = pop.nwalkers
        sw, _ = pop.propagate(rng.uniform(size=n * pop.nC))
        pop.energy(step, True)
        pop.popcontrol(walkers / sw, rng.uniform(size=n), 3 * walkers)
        ws += n
    dt = time.perf_counter() - t0
    del pop
    rate = ws / dt
    return dict(value=rate, unit="walker-steps/s", cores=threads, kind="port", physical_cores=phys,
                gflops=rate * F_REF_PER_WS / 1e9,
                sample=f"{ws} walker-steps ({steps} Walk steps of {walkers} walkers after one warm-up step) of the "
                       f"L={system.L}, D={int(vd.max())} chain: reference algorithm, FP64, {threads} OpenMP "
                       f"threads, {dt:.1f} s")


def learned_comparison(eng, args, first, plain_vt):
    """The learned proposal on the same walk: collect (feature, exact label)
    rows from plain steps, fit the MLP on the GPU, then time as many steps with
    the learned proposal and compare variance x GPU-s with the plain run."""
    from mpsqmc_amd.analysis import variance_time
    from mpsqmc_amd.mlp import MLP
    rows_per_step = args.walkers * eng.nC
    steps = args.var_steps
    n_collect = -(-args.learned_samples // rows_per_step) + 1
    eng.set_timing(False)
    eng.collect_samples(args.learned_samples)
    eng.run(first, n_collect)
    eng.sync()
    n_rows = eng.num_samples()
    net = MLP(seed=args.seed)
    t0 = time.perf_counter()
    loss = eng.train_proposal(net, epochs=args.learned_epochs, batch=4096, lr=1e-3)
    eng.sync()
    t_train = time.perf_counter() - t0
    log(f"bench: learned proposal trained on {n_rows} rows in {t_train:.2f} s (loss {loss:.4g})")
    eng.collect_samples(0)
    eng.set_proposal(net, mix=args.learned_mix)
    first2 = first + n_collect
    t0 = time.perf_counter()
    eng.run(first2, steps)
    eng.sync()
    wall = time.perf_counter() - t0
    tr = eng.read_trace(first2, steps)
    eng.set_proposal(None)
    net.close()
    vt = variance_time(tr[:, 2], wall)
    return {"samples": int(n_rows), "epochs": args.learned_epochs, "train_loss": loss, "train_s": t_train,
            "mix": args.learned_mix, "value": float(tr[:, 1].sum()) / wall, "unit": "walker-steps/s",
            "population_mean": float(tr[:, 1].mean()),
            "variance_x_gpu_s": vt["variance_x_seconds"], "E_mean": vt["mean"], "E_error": vt["error"],
            "block_length": vt["block_length"], "steps": steps, "seconds": wall,
            "gain_vs_plain": (plain_vt["variance_x_seconds"] / vt["variance_x_seconds"]) if plain_vt else None,
            "gain_rel_error": float(np.hypot(plain_vt["var_rel_error"], vt["var_rel_error"])) if plain_vt else None}


def snake_workload(args):
    """Walker-steps/s of one snake lattice on the general path.  4x4: the trial
    from the built-in DMRG (D_T = 64) and walkers set up as SetupWalkers does
    (MPSQMC2.cpp:400-409): copies of the trial, compressed on the device to their
    D_W when D_W < D_T, overlaps set.  6x6: a random left-normalised trial of
    D = 128 (the host DMRG at that size takes longer than the benchmark) and
    walkers that are its copies (D_W = D_T).  Device RNG, population control."""
    from mpsqmc_amd import mps as M
    from mpsqmc_amd import systems
    from mpsqmc_amd.engine import Engine

    w = SNAKES[args.workload]
    system = systems.snake_lattice(w["rows"], w["cols"])
    L, N = system.L, w["N"]
    t0 = time.perf_counter()
    if w["DT"] <= 64 and not args.random_trial:
        tvd, tdata, E_trial = M.dmrg(system, w["DT"], max_sweeps=4, tol=1e-8, seed=args.seed)
    else:
        tvd, tdata = M.random_mps(L, w["DT"], 2, 1)
        tdata, _ = M.left_normalize(tvd, tdata)
        E_trial = None
    log(f"bench: {args.workload} trial ready ({'random' if E_trial is None else f'DMRG, E = {E_trial:.10f}'}, "
        f"{time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(2)
    dws = rng.choice(w["DW"], size=N)
    counts = {int(D): int(np.sum(dws == D)) for D in w["DW"]}
    eng = Engine(system, tvd, tdata, n_desired=N, capacity=3 * N, max_bond=max(w["DW"]), precision=args.precision,
                 seed=args.seed)
    eng.load_copies(tvd, tdata, N)
    targets = np.array([int(D) if D < w["DT"] else 0 for D in dws], dtype=np.int32)
    t0 = time.perf_counter()
    if np.any(targets > 0):
        eng.compress_walkers(targets)
    t_comp = time.perf_counter() - t0
    eng.prepare()
    log(f"bench: {args.workload} engine ready ({N} walkers, L={L}, path {eng.path}; device recompression "
        f"{t_comp:.2f} s); warmup {args.warmup} steps")
    eng.run(1, args.warmup)
    eng.sync()
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tiles, ms0, ms1 = eng.autotune(reps=2)
        tune = {"tiles_by_bond_bucket": [int(t) for t in tiles], "walker_kernel_ms_default": ms0,
                "walker_kernel_ms_tuned": ms1, "gain": ms0 / ms1 if ms1 > 0 else None,
                "seconds": time.perf_counter() - t0}
        log(f"bench: autotuned tiles {list(tiles)}: {ms0:.2f} -> {ms1:.2f} ms per propagation")
    eng.set_timing(True)
    first = args.warmup + 1
    t0 = time.perf_counter()
    for c in range(0, args.steps, 16):  # progress lines: a D = 128 step takes over a second
        eng.run(first + c, min(16, args.steps - c))
        eng.sync()
        log(f"bench: {min(c + 16, args.steps)} steps, {time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    ktot_ms, nlaunch = eng.kernel_time()
    trace = eng.read_trace(first, args.steps)
    eng.close()
    walker_steps = float(trace[:, 1].sum())
    value = walker_steps / wall
    # flops per walker-step of the initial bond mixture (population control
    # copies walkers by weight, which can shift the mixture a little)
    f32 = sum(counts[D] * w["F32"][D] for D in counts) / N
    f64 = sum(counts[D] * w["F64"][D] for D in counts) / N
    avg_launch_s = ktot_ms / 1000.0 / max(nlaunch, 1)
    walkers_per_launch = walker_steps / max(nlaunch, 1)
    achieved = (f32 + f64) * walkers_per_launch / avg_launch_s / 1e12
    peak = (f32 + f64) / (f32 / PEAK_FP32 + f64 / PEAK_FP64) / 1e12
    dws_s = "{" + ",".join(str(D) for D in w["DW"]) + "}"
    out = {
        "metric": f"walker-steps/s ({w['rows']}x{w['cols']} snake, D_T={w['DT']}, D_W in {dws_s}, {N} walkers)",
        "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1000.0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == "fp32" else "f64",
        "data": ("synthetic: trial from the built-in two-site DMRG; walkers = trial copies compressed on the device "
                 "to their D_W" if E_trial is not None else
                 "synthetic: random left-normalized trial; walkers = its copies"),
        "config": {"workload": f"heisenberg {w['rows']}x{w['cols']} OBC snake (L={L}), trial D={w['DT']}, "
                               f"walkers D in {dws_s}, N={N}, dtau=0.01, Marsaglia(4) grid, "
                               f"{args.precision} contractions", "walkers": N, "L": L,
                   "walkers_per_bond": counts, "recompression_s": t_comp},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": "walker_kernel", "avg_launch_ms": avg_launch_s * 1000.0,
                     "flops_per_walker_step": f32 + f64},
        "autotune": tune,
        "E_trial": E_trial,
        "E_P_last": float(trace[-1, 2]),
        "population_mean": float(trace[:, 1].mean()),
        "population_last": int(trace[-1, 1]),
    }
    out["cpu_baseline"] = None
    if not args.no_cpu and w["DT"] <= 64:
        out["cpu_baseline"] = snake_cpu_baseline(system, tvd, tdata, w, args.cpu_threads)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def snake_cpu_baseline(system, tvd, tdata, w, threads):
    """The reference algorithm on the snake lattice: one Walk step (after a
    warm-up step) of 4 walkers per D_W, each the trial compressed to its D_W."""
    import oracle
    use, phys, aff = physical_cores()
    threads = min(threads or use, 16, aff)
    L = system.L
    vds, datas, Os = [], [], []
    for D in w["DW"]:
        if D < w["DT"]:
            vd, data, _ = oracle.compress(L, 2, tvd, tdata, D)
            _, data = oracle.left_normalize(L, 2, vd, data)
            O = oracle.inner(L, 2, tvd, tdata, vd, data)
        else:
            vd, data, O = tvd, tdata, 1.0
        vds += [vd] * 4
        datas += [data] * 4
        Os += [abs(O)] * 4
    n = len(vds)
    pop = oracle.Population(L, 2, tvd, tdata, system.J, system.h, system.dtau, system.tables_distinct(),
                            system.grid().reshape(-1), 4, build_terms=True, threads=threads, energy_mode=1)
    rng = np.random.default_rng(1)
    pop.set_walkers(vds[:threads], datas[:threads], np.ones(min(n, threads)), np.array(Os[:threads]),
                    np.zeros(min(n, threads)))
    pop.propagate(rng.uniform(size=pop.nwalkers * pop.nC))  # warm-up step
    pop.set_walkers(vds, datas, np.ones(n), np.array(Os), np.zeros(n))
    t0 = time.perf_counter()
    sw, _ = pop.propagate(rng.uniform(size=n * pop.nC))
    pop.energy(1, True)
    pop.popcontrol(n / sw, rng.uniform(size=n), 3 * n)
    dt = time.perf_counter() - t0
    del pop
    return dict(value=n / dt, unit="walker-steps/s", cores=threads, kind="port", physical_cores=phys,
                sample=f"{n} walker-steps (one Walk step of 4 walkers per D_W after a warm-up step) of the "
                       f"{system.L}-site snake: reference propagation (16 + 1 full overlaps per coupling), local "
                       f"energy by operator insertion, FP64, {threads} OpenMP threads, {dt:.1f} s")


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(walker sharding over GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "chain":
        print(json.dumps(snake_workload(args)), flush=True)
        return
    from mpsqmc_amd import systems
    from mpsqmc_amd.engine import Engine

    L, D, N = args.L, args.D, args.walkers
    system = systems.chain(L, J=1.0, dtau=0.01)
    t0 = time.perf_counter()
    vd, data, E_trial = make_trial(system, D, args.seed, args.random_trial)
    log(f"bench: trial ready ({'random' if E_trial is None else f'DMRG, E = {E_trial:.10f}'}, "
        f"{time.perf_counter() - t0:.1f} s)")
    eng = Engine(system, vd, data, n_desired=N, capacity=3 * N, max_bond=D, precision=args.precision,
                 seed=args.seed)
    eng.load_copies(vd, data, N)
    eng.prepare()
    log(f"bench: engine ready ({N} walkers, L={L}, D={D}, {args.precision}); warmup {args.warmup} steps")
    eng.run(1, args.warmup)
    eng.sync()
    log("bench: warmup done; timing")
    eng.set_timing(True)
    first = args.warmup + 1
    t0 = time.perf_counter()
    eng.run(first, args.steps)
    eng.sync()
    t1 = time.perf_counter()
    log(f"bench: {args.steps} steps in {t1 - t0:.3f} s")
    ktot_ms, nlaunch = eng.kernel_time()
    trace = eng.read_trace(first, args.steps)
    walker_steps = float(trace[:, 1].sum())
    wall = t1 - t0
    value = walker_steps / wall
    from mpsqmc_amd.analysis import variance_time
    # roofline of the dominant kernel (the fused walker kernel)
    avg_launch_s = ktot_ms / 1000.0 / max(nlaunch, 1)
    walkers_per_launch = walker_steps / max(nlaunch, 1)
    flops_per_launch = (F32_PER_WS + F64_PER_WS) * walkers_per_launch
    achieved = flops_per_launch / avg_launch_s / 1e12
    t_min = F32_PER_WS / PEAK_FP32 + F64_PER_WS / PEAK_FP64
    peak = (F32_PER_WS + F64_PER_WS) / t_min / 1e12
    out = {
        "metric": "walker-steps/s (L=32, D=32, 2048 walkers)",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "fp32" else "f64",
        "data": ("synthetic: random left-normalized trial MPS with the D bond profile" if E_trial is None else
                 "synthetic: trial from the built-in two-site DMRG of the chain (D bond profile)") +
                "; walkers start as copies of the trial",
        "config": {"workload": f"heisenberg-chain L={L} OBC J=1, trial D={D}, walkers D={D}, N={N}, dtau=0.01, "
                               f"Marsaglia(4) grid, {args.precision} contractions",
                   "walkers": N, "L": L, "D": D},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(L, D, N, walkers_per_launch),
                     "traffic_source": "profiles/walker32_kernel_chain_L32_D32_N2048_hbm.md (rocprofv3 FETCH_SIZE "
                                       "+ WRITE_SIZE per walker-step, scaled by a copy kernel) x walkers per launch",
                     "kernel": eng.kernel_name, "avg_launch_ms": avg_launch_s * 1000.0,
                     "flops_per_walker_step": F32_PER_WS + F64_PER_WS},
        "E_trial": E_trial,
        "E_P_last": float(trace[-1, 2]),
        "population_mean": float(trace[:, 1].mean()),
        "population_last": int(trace[-1, 1]),
    }
    nxt = first + args.steps
    eng.set_timing(False)
    if not args.no_graph:
        # the same steps replayed from a captured graph (mqe_run with use_graph)
        eng.run(nxt, args.steps, use_graph=True)  # capture and a first replay
        eng.sync()
        nxt += args.steps
        t0 = time.perf_counter()
        eng.run(nxt, args.steps, use_graph=True)
        eng.sync()
        gw = time.perf_counter() - t0
        tg = eng.read_trace(nxt, args.steps)
        nxt += args.steps
        out["graph_replay"] = {"value": float(tg[:, 1].sum()) / gw, "unit": "walker-steps/s",
                               "ms_per_step": gw / args.steps * 1000.0, "steps": args.steps}
        log(f"bench: graph replay {args.steps} steps in {gw:.3f} s")
    # metric 2: variance of the mean projected energy x GPU seconds, on a fixed
    # number of steps whatever --steps is
    vt = None
    if args.var_steps >= 64:
        t0 = time.perf_counter()
        eng.run(nxt, args.var_steps)
        eng.sync()
        vw = time.perf_counter() - t0
        tv = eng.read_trace(nxt, args.var_steps)
        nxt += args.var_steps
        vt = variance_time(tv[:, 2], vw)
        out["variance_x_gpu_s"] = {
            "value": vt["variance_x_seconds"], "unit": "E^2*s", "proposal": "plain",
            "E_mean": vt["mean"], "E_error": vt["error"], "block_length": vt["block_length"],
            "steps": vt["n"], "seconds": vt["seconds"], "walker_steps_per_s": float(tv[:, 1].sum()) / vw}
        log(f"bench: variance run {args.var_steps} steps in {vw:.2f} s")
        if not args.no_learned and eng.path == "tile32":
            out["learned_proposal"] = learned_comparison(eng, args, nxt, vt)
    else:
        out["variance_x_gpu_s"] = None
    eng.close()
    if not args.no_cpu:
        log(f"bench: CPU baseline on {args.cpu_walkers} walkers")
        out["cpu_baseline"] = cpu_baseline(system, vd, data, args.cpu_walkers, args.cpu_steps, args.cpu_threads)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
