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
variance of the mean projected energy over the timed steps times their wall
time (the cost of a run at a fixed statistical error is proportional to it).

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
    ap.add_argument("--cpu-walkers", type=int, default=512)
    ap.add_argument("--cpu-threads", type=int, default=16)
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


def cpu_baseline(system, vd, data, walkers, threads):
    """The reference algorithm (16 + 1 full-chain overlaps per coupling, a full
    QR left-normalisation after every operator, energy through H|trial>) in the
    FP64 C restatement, OpenMP over walkers, on a bounded sample."""
    import oracle
    threads = max(1, min(threads, os.cpu_count() or 1))
    pop = oracle.Population(system.L, 2, vd, data, system.J, system.h, system.dtau, system.tables_distinct(),
                            system.grid().reshape(-1), 4, build_terms=True, threads=threads, energy_mode=2)
    pop.init_walkers(walkers, int(vd.max()))
    rng = np.random.default_rng(1)
    u = rng.uniform(size=walkers * pop.nC)
    t0 = time.perf_counter()
    pop.propagate(u)
    pop.energy(1, True)
    dt = time.perf_counter() - t0
    del pop
    return dict(value=walkers / dt, unit="walker-steps/s", cores=threads, kind="port",
                sample=f"{walkers} walkers x 1 step of the L={system.L}, D={int(vd.max())} chain "
                       f"(reference algorithm, FP64, {threads} OpenMP threads), {dt:.1f} s")


def learned_comparison(eng, args, first, plain_vt):
    """The learned proposal on the same walk: collect (feature, exact label)
    rows from plain steps, fit the MLP on the GPU, then time as many steps with
    the learned proposal and compare variance x GPU-s with the plain run."""
    from mpsqmc_amd.analysis import variance_time
    from mpsqmc_amd.mlp import MLP
    rows_per_step = args.walkers * eng.nC
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
    eng.run(first2, args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    tr = eng.read_trace(first2, args.steps)
    eng.set_proposal(None)
    net.close()
    vt = variance_time(tr[:, 2], wall)
    return {"samples": int(n_rows), "epochs": args.learned_epochs, "train_loss": loss, "train_s": t_train,
            "mix": args.learned_mix, "value": float(tr[:, 1].sum()) / wall, "unit": "walker-steps/s",
            "population_mean": float(tr[:, 1].mean()),
            "variance_x_gpu_s": vt["variance_x_seconds"], "E_mean": vt["mean"], "E_error": vt["error"],
            "block_length": vt["block_length"], "steps": args.steps, "seconds": wall,
            "gain_vs_plain": (plain_vt["variance_x_seconds"] / vt["variance_x_seconds"]) if plain_vt else None,
            "gain_rel_error": float(np.hypot(plain_vt["var_rel_error"], vt["var_rel_error"])) if plain_vt else None}


def snake_workload(args):
    """Walker-steps/s of one snake lattice on the general path: random
    left-normalised trial and walker MPS of the stated bonds (throughput does
    not depend on the tensor values), device RNG, population control."""
    from mpsqmc_amd import mps as M
    from mpsqmc_amd import systems
    from mpsqmc_amd.engine import Engine

    w = SNAKES[args.workload]
    system = systems.snake_lattice(w["rows"], w["cols"])
    L, N = system.L, w["N"]
    tvd, tdata = M.random_mps(L, w["DT"], 2, 1)
    tdata, _ = M.left_normalize(tvd, tdata)
    rng = np.random.default_rng(2)
    states, vds, datas, counts = {}, [], [], {D: 0 for D in w["DW"]}
    for _ in range(N):
        D = int(rng.choice(w["DW"]))
        if D not in states:
            vd, data = M.random_mps(L, D, 2, 10 + D)
            states[D] = (vd, M.left_normalize(vd, data)[0])
        vds.append(states[D][0])
        datas.append(states[D][1])
        counts[D] += 1
    eng = Engine(system, tvd, tdata, n_desired=N, capacity=3 * N, max_bond=max(w["DW"]), precision=args.precision,
                 seed=args.seed)
    eng.load_walkers(vds, datas)
    eng.prepare()
    log(f"bench: {args.workload} engine ready ({N} walkers, L={L}, path {eng.path}); warmup {args.warmup} steps")
    eng.run(1, args.warmup)
    eng.sync()
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
    dws = "{" + ",".join(str(D) for D in w["DW"]) + "}"
    return {
        "metric": f"walker-steps/s ({w['rows']}x{w['cols']} snake, D_T={w['DT']}, D_W in {dws}, {N} walkers)",
        "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1000.0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == "fp32" else "f64",
        "data": "synthetic: random left-normalized trial and walker MPS of the stated bond profiles",
        "config": {"workload": f"heisenberg {w['rows']}x{w['cols']} OBC snake (L={L}), trial D={w['DT']}, "
                               f"walkers D in {dws}, N={N}, dtau=0.01, Marsaglia(4) grid, "
                               f"{args.precision} contractions", "walkers": N, "L": L},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": "walker_kernel", "avg_launch_ms": avg_launch_s * 1000.0,
                     "flops_per_walker_step": f32 + f64},
        "cpu_baseline": None,
        "E_P_last": float(trace[-1, 2]),
        "population_mean": float(trace[:, 1].mean()),
        "population_last": int(trace[-1, 1]),
    }


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
    if args.steps % 64:
        log("bench: note: the tile path re-orthonormalises every 64 steps; time a multiple of 64 steps "
            "for the amortised rate")
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
    vt = variance_time(trace[:, 2], wall) if args.steps >= 64 else None
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
                     "kernel": eng.kernel_name, "avg_launch_ms": avg_launch_s * 1000.0,
                     "flops_per_walker_step": F32_PER_WS + F64_PER_WS},
        "variance_x_gpu_s": None if vt is None else {
            "value": vt["variance_x_seconds"], "unit": "E^2*s", "proposal": "plain",
            "E_mean": vt["mean"], "E_error": vt["error"], "block_length": vt["block_length"],
            "steps": vt["n"], "seconds": vt["seconds"]},
        "E_trial": E_trial,
        "E_P_last": float(trace[-1, 2]),
        "population_mean": float(trace[:, 1].mean()),
        "population_last": int(trace[-1, 1]),
    }
    if not args.no_learned and vt is not None and eng.path == "tile32":
        out["learned_proposal"] = learned_comparison(eng, args, first + args.steps, vt)
    eng.close()
    if not args.no_cpu:
        log(f"bench: CPU baseline on {args.cpu_walkers} walkers")
        out["cpu_baseline"] = cpu_baseline(system, vd, data, args.cpu_walkers, args.cpu_threads)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
