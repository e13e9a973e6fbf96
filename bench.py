#!/usr/bin/env python3
"""Benchmark: Burgers (Parthenon-VIBE) RK2 cycle throughput on one MI355X.

Default workload (--workload burgers-amr), the static-refined Burgers step: a 64^3 base mesh on
[-0.5, 0.5]^3 in 16^3 meshblocks, refined two levels on every block that meets the sphere of
radius 0.25 at the centre (960 blocks with 2:1 nesting), nghost 4, 3 velocities + 4 passive
scalars, FP64, WENO5 + HLL, RK2, CFL 0.4, periodic, seeded sum-of-sinusoids initial state
(SplitMix64 seed 2). --workload burgers-uniform is the single-level 64^3 mesh in 64 blocks
(seed 0). One step = one full RK2 cycle: per stage the reconstruction, the flux-correction
launch (multilevel), the fused HLL/flux-divergence/RK-update/derived launch and the boundary
exchange (restriction, direct same-GPU copy over the buffer table, restriction of received
regions, minmod prolongation); the CFL estimate and SetGlobalTimeStep on device at the end.
Inputs are resident in HBM before timing starts. Throughput is zone-cycles/s
= (sum of block cells) * steps / wall time (the reference's definition,
src/driver/driver.cpp:56-63).

Timed region 1 (the reported value): `steps` cycles, each one replay of the captured hipGraph
(--driver graph, default), bracketed by stream synchronizations. Timed region 2 (roofline):
the same cycles launched eagerly with HIP events around every kernel on its stream; the
dominant kernel's average duration and its algorithmic bytes give `roofline.achieved`.

The CPU baseline is the oracle's C++ restatement of the reference (OpenMP over meshblocks)
on a bounded sample of the same workload, run in a child process.
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


def kernel_bytes(nb, n, g, V, exch_elems):
    """Algorithmic HBM bytes per launch of each kernel (FP64; the compulsory unique traffic of
    the kernel's inputs and outputs)."""
    E = (n + 2 * g) ** 3
    I = n ** 3
    F = (n + 1) * n * n
    RU = I + 18 * n * n
    recon = 8 * V * nb * (RU + 6 * (n + 2) * n * n)
    return {
        "recon": recon,
        "recon_tasks": recon,
        "fused_update": 8 * nb * (6 * V * F + 3 * V * I + I),
        "hll_flux": 8 * V * nb * 9 * F,
        "flux_div": 8 * V * nb * (3 * F + I),
        "weighted_sum": 8 * V * nb * 3 * E,
        "derived": 8 * nb * 5 * I,
        "dt_reduce": 8 * nb * 3 * I,
        "pack": 16 * exch_elems,
        "unpack": 16 * exch_elems,
        "exchange_direct": 16 * exch_elems,
    }


def step_bytes(nb, n, g, V, exchange_bytes):
    """Algorithmic bytes of one RK2 cycle in the reference task decomposition: per block the
    reconstruction/flux, divergence, averaging, update, derived and dt tasks, plus the
    exchange's buffer traffic."""
    E = (n + 2 * g) ** 3
    I = n ** 3
    F = (n + 1) * n * n
    RU = I + 18 * n * n
    return 8 * nb * (V * (2 * RU + 12 * F + 2 * I + 11 * E) + 13 * I) + exchange_bytes


def cpu_baseline(nx, block, g, ns, spheres, seed, cycles, threads):
    """Oracle (C++ restatement, OpenMP over meshblocks) on `cycles` cycles of the workload,
    run in a child process so its OpenMP thread count is set before the library loads."""
    code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import OracleMesh, OracleBurgers, lib
m = OracleMesh({nx}, {block}, {g}, spheres={list(spheres)!r})
b = OracleBurgers(m, {ns}, 0.4, 0)
lib().orc_burgers_pgen_sinusoid(b.h, {seed})
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


def copy_bench(args):
    """Plain device copy of 1 GiB (one read and one write per byte): the reference kernel for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes."""
    import torch
    n = 1 << 27
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nbytes = 2 * 8 * n
    print(json.dumps({"metric": "device copy bandwidth (counter calibration)",
                      "value": nbytes * args.steps / wall / 1e9, "unit": "GB/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n,
                      "higher_is_better": True, "dtype": "f64", "data": "synthetic"}))


def exchange_bench(args):
    """Ghost-exchange microbenchmark: 128^3 mesh in 512 blocks of 16^3, nghost 3, 8 FP64
    one-component fields, U(-1,1) interiors; one step = SendBoundBufs + ReceiveBoundBufs +
    SetBounds over all 26 neighbour buffers of every block."""
    import parthenon_amd as pa
    from parthenon_amd.inputs import interior_uniform

    m = pa.Mesh((128, 128, 128), (16, 16, 16), 3)
    for f in range(8):
        m.add_field(f"q{f}", 1)
    m.finalize()
    m.set_pack_mode(pa.PackMode.per_buffer if args.pack == "per_buffer" else pa.PackMode.persistent)
    md = m.mesh_data("base")
    for f in range(8):
        md.upload(f"q{f}", interior_uniform(m.field_shape(f"q{f}"), m.ng, seed=1000 + f, fill=0.0))
    xchg = pa.exchange_fused if args.pack == "fused" else pa.exchange
    for _ in range(args.warmup):
        xchg(md)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        xchg(md)
    m.synchronize()
    wall = time.perf_counter() - t0
    pa.profile_enable(True)
    pa.profile_report()
    for _ in range(min(args.steps, 50)):
        xchg(md)
    m.synchronize()
    prof = pa.profile_report()
    pa.profile_enable(False)
    nbuf, nel = m.exchange_counts()
    alg = 4 * 8 * nel  # pack R+W and unpack R+W, FP64
    us = {k: ms / c * 1e3 for k, (c, ms) in prof.items()}
    dom = max(prof, key=lambda k: prof[k][1])
    achieved = 16 * nel / (us[dom] * 1e-6) / 1e9
    print(json.dumps({
        "metric": "boundary exchanges/s (Send+Recv+Set, 26 buffers per block)",
        "value": args.steps / wall, "unit": "exchanges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic U(-1,1)",
        "config": {"workload": "exchange 128^3 mesh, 512 blocks of 16^3, nghost 3, 8 fields",
                   "pack": args.pack, "buffers": nbuf, "elements": nel},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": 16 * nel, "avg_launch_us": us[dom]},
        "exchange_roofline": {"algorithmic_bytes": alg,
                              "achieved_GBps": alg / (wall / args.steps) / 1e9},
        "kernels_avg_us": us}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload",
                    choices=["burgers-amr", "burgers-uniform", "exchange", "copy"],
                    default="burgers-amr")
    ap.add_argument("--pack", choices=["persistent", "per_buffer", "fused"], default="persistent")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--driver", choices=["graph", "eager", "tasks"], default="graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "exchange":
        return exchange_bench(args)
    if args.workload == "copy":
        return copy_bench(args)

    import parthenon_amd as pa
    from parthenon_amd.burgers import BurgersDriver, DeviceBurgersDriver, make_burgers_mesh

    nx, block, g, ns = (64, 64, 64), (16, 16, 16), 4, 4
    amr = args.workload == "burgers-amr"
    spheres = [((0.0, 0.0, 0.0), 0.25, 2)] if amr else []
    seed = 2 if amr else 0
    cpu_cycles = args.cpu_cycles or (10 if amr else 60)
    V = 3 + ns
    mesh = make_burgers_mesh(nx, block, g, ns, refine_spheres=spheres)
    if args.driver == "tasks":
        drv = BurgersDriver(mesh, cfl=0.4)
        drv.initialize("sinusoid", seed=seed)
        run = lambda k: [drv.cycle() for _ in range(k)]
        sync = mesh.synchronize
    else:
        drv = DeviceBurgersDriver(mesh, cfl=0.4, use_graph=(args.driver == "graph"))
        drv.initialize("sinusoid", seed=seed)
        run = drv.cycles
        sync = drv.synchronize

    run(args.warmup)
    sync()
    t0 = time.perf_counter()
    run(args.steps)
    sync()
    wall = time.perf_counter() - t0

    # roofline pass: the same cycles launched eagerly, HIP events around every kernel
    prof_steps = min(args.steps, 50 if not amr else 10)
    if args.driver == "graph":
        drv.use_graph = False
    pa.profile_enable(True)
    pa.profile_report()
    run(prof_steps)
    sync()
    prof = pa.profile_report()
    pa.profile_enable(False)

    nb = mesh.nblocks
    n = block[0]
    value = nb * n ** 3 * args.steps / wall
    _, nel = mesh.exchange_counts()
    kb = kernel_bytes(nb, n, g, V, nel)
    prof.pop("graph_cycle", None)
    dom = max(prof, key=lambda k: prof[k][1])
    launches, total_ms = prof[dom]
    avg_s = total_ms / launches * 1e-3
    achieved = kb[dom] / avg_s / 1e9 if dom in kb else None
    kernels = {k: {"launches_per_step": c / prof_steps, "avg_us": ms / c * 1e3,
                   "GBps": (kb[k] / (ms / c * 1e-3) / 1e9) if k in kb else None}
               for k, (c, ms) in sorted(prof.items())}
    # interior part of the reference task decomposition plus the exchange's buffer traffic
    # counted from the live table (send R+W and set R+W per element, two stages); restriction,
    # prolongation and flux-correction traffic is not counted (a lower bound)
    sb = step_bytes(nb, n, g, V, 2 * 4 * 8 * nel)
    if amr:
        workload = ("burgers-amr 64^3 base mesh in 16^3 blocks, two static levels on the blocks "
                    "meeting the centred sphere r=0.25 (960 blocks), nghost 4, 3 velocities + "
                    "4 scalars, WENO5/HLL/RK2 with flux correction, CFL 0.4, periodic")
    else:
        workload = ("burgers-uniform 64^3 mesh, 64 blocks of 16^3, nghost 4, 3 velocities + "
                    "4 scalars, WENO5/HLL/RK2, CFL 0.4, periodic")
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
        "data": f"synthetic (seeded sum of sinusoids, SplitMix64 seed {seed})",
        "config": {"workload": workload, "driver": args.driver, "nblocks": nb,
                   "block": list(block), "nghost": g, "ncomp": V},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": None, "algorithmic_bytes_per_launch": kb.get(dom),
                     "avg_launch_us": avg_s * 1e6},
        "step_roofline": {"algorithmic_bytes_per_step": sb,
                          "achieved_GBps": sb / (wall / args.steps) / 1e9,
                          "frac": sb / (wall / args.steps) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        threads = min(len(os.sched_getaffinity(0)), 16)
        cw = cpu_baseline(nx, block, g, ns, spheres, seed, cpu_cycles, threads)
        out["cpu_baseline"] = {
            "value": nb * n ** 3 * cpu_cycles / cw, "unit": "zone-cycles/s",
            "cores": threads, "kind": "port",
            "sample": f"{cpu_cycles} RK2 cycles of the same workload after 1 warm-up cycle, "
                      f"oracle C++ restatement with OpenMP over meshblocks ({cw:.2f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
