#!/usr/bin/env python3
"""Benchmark: Burgers (Parthenon-VIBE) RK2 cycle throughput on one MI355X.

Workload (default): 64^3 mesh in 64 meshblocks of 16^3, nghost 4, 3 velocities + 4 passive
scalars, FP64, WENO5 + HLL, RK2, CFL 0.4, periodic, seeded sum-of-sinusoids initial state
(SplitMix64 seed 0). One step = one full RK2 cycle (two stages: fluxes, flux divergence,
averaging, update, boundary exchange, derived field; time-step estimate on the last stage).
Inputs are resident in HBM before timing starts. Throughput is zone-cycles/s
= nblocks * 16^3 * steps / wall time (the reference's definition, src/driver/driver.cpp:56-63).

Prints ONE JSON line with the metric, per-kernel roofline of the dominant kernel (HIP events on
the launch stream over the timed region), and the CPU baseline: the oracle's C++ restatement of
the reference (OpenMP over meshblocks) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(nb, n, g, V):
    """Algorithmic HBM bytes per launch of each stage kernel (FP64, compulsory unique traffic
    of the task's inputs and outputs)."""
    E = (n + 2 * g) ** 3
    I = n ** 3
    F = (n + 1) * n * n
    RU = I + 18 * n * n
    return {
        "recon": 8 * V * nb * (RU + 6 * (n + 2) * n * n),
        "hll_flux": 8 * V * nb * 9 * F,
        "flux_div": 8 * V * nb * (3 * F + I),
        "weighted_sum": 8 * V * nb * 3 * E,
        "derived": 8 * nb * 5 * I,
        "dt_reduce": 8 * nb * 3 * I,
    }


def step_bytes(nb, n, g, V):
    """Algorithmic bytes of one RK2 cycle in the reference task decomposition."""
    E = (n + 2 * g) ** 3
    I = n ** 3
    F = (n + 1) * n * n
    RU = I + 18 * n * n
    H = E - I
    return 8 * nb * (V * (2 * RU + 12 * F + 2 * I + 11 * E + 8 * H) + 13 * I)


def cpu_baseline(nx, block, g, ns, cycles, threads):
    """Oracle (C++ restatement, OpenMP over meshblocks) on `cycles` cycles of the workload,
    run in a child process so its OpenMP thread count is set before the library loads."""
    code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import OracleMesh, OracleBurgers, lib
m = OracleMesh({nx}, {block}, {g})
b = OracleBurgers(m, {ns}, 0.4, 0)
lib().orc_burgers_pgen_sinusoid(b.h, 0)
lib().orc_burgers_initialize(b.h)
lib().orc_burgers_cycles(b.h, 1)
t = time.perf_counter()
lib().orc_burgers_cycles(b.h, {cycles})
dt = time.perf_counter() - t
print(json.dumps({{"wall": dt}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])["wall"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=60)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import parthenon_amd as pa
    from parthenon_amd.burgers import BurgersDriver, make_burgers_mesh

    nx, block, g, ns = (64, 64, 64), (16, 16, 16), 4, 4
    V = 3 + ns
    mesh = make_burgers_mesh(nx, block, g, ns)
    drv = BurgersDriver(mesh, cfl=0.4)
    drv.initialize("sinusoid", seed=0)
    for _ in range(args.warmup):
        drv.cycle()
    mesh.synchronize()

    pa.profile_enable(True)
    pa.profile_report()  # drop warm-up records
    t0 = time.perf_counter()
    for _ in range(args.steps):
        drv.cycle()
    mesh.synchronize()
    wall = time.perf_counter() - t0
    prof = pa.profile_report()
    pa.profile_enable(False)

    nb = mesh.nblocks
    n = block[0]
    zone_cycles = nb * n ** 3 * args.steps
    value = zone_cycles / wall
    kb = kernel_bytes(nb, n, g, V)
    dom = max(prof, key=lambda k: prof[k][1])
    launches, total_ms = prof[dom]
    avg_s = total_ms / launches * 1e-3
    achieved = kb.get(dom, 0) / avg_s / 1e9 if dom in kb else None
    kernels = {k: {"launches": c, "avg_us": ms / c * 1e3,
                   "GBps": (kb[k] / (ms / c * 1e-3) / 1e9) if k in kb else None}
               for k, (c, ms) in prof.items()}
    out = {
        "metric": "zone-cycles/s (Burgers RK2 cycle, FP64)",
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sum of sinusoids, SplitMix64 seed 0)",
        "config": {"workload": "burgers-vibe 64^3 mesh, 64 blocks of 16^3, nghost 4, "
                               "3 velocities + 4 scalars, WENO5/HLL/RK2, CFL 0.4, periodic",
                   "nblocks": nb, "block": list(block), "nghost": g, "ncomp": V},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": None,
                     "algorithmic_bytes_per_launch": kb.get(dom)},
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes(nb, n, g, V),
                          "achieved_GBps": step_bytes(nb, n, g, V) / (wall / args.steps) / 1e9},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        threads = min(len(os.sched_getaffinity(0)), 16)
        cw = cpu_baseline(nx, block, g, ns, args.cpu_cycles, threads)
        out["cpu_baseline"] = {
            "value": nb * n ** 3 * args.cpu_cycles / cw, "unit": "zone-cycles/s",
            "cores": threads, "kind": "port",
            "sample": f"{args.cpu_cycles} RK2 cycles of the same workload after 1 warm-up cycle, "
                      f"oracle C++ restatement with OpenMP over meshblocks ({cw:.2f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
