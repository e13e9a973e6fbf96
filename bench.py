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


def same_level_ghost_cells(mesh):
    """Ghost cells of every block whose neighbour across them is a block of the same level (a
    face slab g x n x n, an edge g x g x n, a corner g^3)."""
    total = 0
    for gid in range(mesh.nblocks):
        level = mesh.block(gid)["level"]
        for nb in mesh.neighbors(gid):
            if nb["level"] != level:
                continue
            cells = 1
            for d, o in enumerate(nb["off"]):
                cells *= mesh.ng[d] if o != 0 else mesh.n[d]
            total += cells
    return total


def kernel_bytes(nb, n, g, V, exch_elems):
    """Algorithmic HBM bytes per launch of each kernel (FP64; the compulsory unique traffic of
    the kernel's inputs and outputs)."""
    E = (n + 2 * g) ** 3
    I = n ** 3
    F = (n + 1) * n * n
    RU = I + 18 * n * n
    recon = 8 * V * nb * (RU + 6 * (n + 2) * n * n)
    # plane kernels: part 0 reads every cell its stencils touch (interior + 3 ghost layers)
    # and writes the 3 face-flux arrays; the ghost-free part reads the interior only and writes
    # the faces between cells 3 .. n - 2 of each row; the boundary part reads the cells within
    # 5 of a block face (plus the ghosts) and writes the 6 faces per row next to the boundary
    inner_faces = max(n - 5, 0) * n * n
    outer_cells = I - max(n - 10, 0) ** 3 + 18 * n * n
    return {
        # block sweep: U over the stencil union and base's interior read, the output and the
        # derived field written (stage 2, where base and U are different arrays)
        "block_sweep": 8 * nb * (V * (RU + 2 * I) + I),
        "recon_flux": 8 * V * nb * (RU + 3 * F),
        "recon_flux_inner": 8 * V * nb * (I + 3 * inner_faces),
        "recon_flux_outer": 8 * V * nb * (outer_cells + 3 * 6 * n * n),
        "flux_update": 8 * nb * (3 * V * F + 3 * V * I + I),
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


def measured_traffic(workload, kernel):
    """HBM-side bytes per launch of `kernel` in `workload` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json, scaled by the copy calibration),
    or None when the summary does not cover that kernel."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(workload)
    if e and "kernels" in e:
        e = e["kernels"].get(kernel)
    elif e and e.get("kernel") != kernel:
        e = None
    if not e:
        return None
    return (e["fetch_kib"] * t["fetch_scale"] + e["write_kib"] * t["write_scale"]) * 1024


def host_cores():
    """Threads for the CPU baseline: every core this process may run on, bounded by the
    OMP_NUM_THREADS the environment grants (16 on a one-GPU share of the box), and the host's
    description for the record."""
    avail = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cpus": avail,
                     "omp_num_threads_env": env or None, "cpu_model": model}


VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4  # wave64 VALU instructions/s: 1024 SIMDs, 4 cycles each


def measured_valu(workload, kernel):
    """Wave-level VALU instructions of `kernel` per launch (for the block sweep: per RK stage,
    summed over the stage's launches) from the committed SQ_INSTS_VALU counter pass
    (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(workload) or {}
    except (OSError, ValueError):
        return None
    return e.get("sq_insts_valu") if e.get("kernel") == kernel else None


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
    """Oracle (C++ restatement, OpenMP over meshblocks): the median of three timings of
    `cycles` consecutive cycles of the workload, run in a child process so its OpenMP thread
    count is set before the library loads."""
    code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import OracleMesh, OracleBurgers, lib
m = OracleMesh({nx}, {block}, {g}, spheres={list(spheres)!r})
b = OracleBurgers(m, {ns}, 0.4, 0)
lib().orc_burgers_pgen_sinusoid(b.h, {seed})
lib().orc_burgers_initialize(b.h)
lib().orc_burgers_cycles(b.h, 1)
walls = []
for _ in range(3):
    t = time.perf_counter()
    lib().orc_burgers_cycles(b.h, {cycles})
    walls.append(time.perf_counter() - t)
print(json.dumps({{"wall": sorted(walls)[1]}}))
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


def gmg_cpu_baseline(N, seed, threads):
    """Oracle FP64 GMG solve of the same problem in a child process (OpenMP loops)."""
    code = f"""
import sys, time, json
import numpy as np
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import OracleGmg
g = OracleGmg({N})
g.rhs_blobs({seed})
t = time.perf_counter()
it, hist = g.solve(tol=1e-10, maxit=100)
dt = time.perf_counter() - t
np.save(sys.argv[1], g.u())
print(json.dumps({{"wall": dt, "iters": int(it)}}))
"""
    import tempfile
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "u.npy")
        r = subprocess.run([sys.executable, "-c", code, path], env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("gmg cpu baseline failed: " + r.stderr[-2000:])
        import numpy as np
        return json.loads(r.stdout.strip().splitlines()[-1]), np.load(path)


def gmg_bench(args, quick=False):
    """Mixed-precision Poisson GMG: 256^3 cells on [0,1]^3 (512 blocks of 32^3 in the
    reference's layout), 7-point operator, FixedFace-0 boundaries, 16 seeded Gaussian blobs
    (SplitMix64 seed 3), V-cycles of 2+2 weighted-Jacobi sweeps (omega 2/3) in FP32 inside
    FP64 iterative refinement, to 1e-10 relative residual. One step = one solve from u = 0.
    The same solves with the FP64 solver are timed beside it."""
    import numpy as np
    import parthenon_amd as pa
    from parthenon_amd.gmg import MGSolver

    from parthenon_amd.gmg import MGParams
    N, seed = 256, 3
    res = {}
    sols = {}
    configs = {"mixed": ("mixed", args.inner_cycles, True), "fp64": ("fp64", 1, True)}
    if not quick:
        if args.inner_cycles != 1:
            configs["mixed_inner1"] = ("mixed", 1, True)
        configs["mixed_cell_sweeps"] = ("mixed", args.inner_cycles, False)
    for name, (prec, inner, cols) in configs.items():
        s = MGSolver(N, prec, MGParams(inner_cycles=inner, column_sweeps=cols))
        s.set_rhs_blobs(seed)
        for _ in range(args.warmup):
            s.reset()
            s.solve()
        steps = args.steps
        t0 = time.perf_counter()
        for _ in range(steps):
            s.reset()
            its, rel = s.solve()
        wall = time.perf_counter() - t0
        # roofline pass: one more solve with HIP events around every launch
        pa.profile_enable(True)
        pa.profile_report()
        s.reset()
        its, rel = s.solve()
        prof = pa.profile_report()
        pa.profile_enable(False)
        res[name] = dict(wall=wall, steps=steps, iters=its, rel=rel, prof=prof)
        sols[name] = s.solution()
        del s
    m = res["mixed"]
    c = N ** 3
    es = 4  # FP32 inner
    # algorithmic bytes per launch of the finest-level kernels
    kb = {"gmg_outer": 32 * c, "gmg_pre_l0": 2 * es * c, "gmg_jacobi_l0": 3 * es * c,
          "gmg_restrict_l0": 2 * es * c + es * c // 8,
          "gmg_prolong_l0": 2 * es * c + es * c // 8,
          "gmg_prolong_jacobi_l0": 3 * es * c + es * c // 8}
    prof = m["prof"]
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][1])
    n_l, ms = prof[dom]
    avg_s = ms / n_l * 1e-3
    achieved = kb[dom] / avg_s / 1e9
    # per outer iteration: the finest level's traffic plus 1/7 of it for the coarser levels;
    # a V-cycle's level traffic: pre-smoothing (b read, x written), residual-restriction (x, b
    # read, coarse b written), prolongation inside the first post-sweep (x, b, coarse x read,
    # t written) and the second post-sweep (x, b read, t written)
    fused = "gmg_prolong_jacobi_l0" in prof
    vc = (2 * es + (2 * es + es / 8) + (3 * es + es / 8) + 3 * es if fused
          else 2 * es + 2 * 3 * es + 2 * (2 * es + es / 8))
    it_bytes = int(32 * c + vc * c * 8 / 7)
    per_it = m["wall"] / m["steps"] / max(m["iters"] * args.inner_cycles, 1)
    out = {
        "metric": "Poisson GMG solves/s (256^3, 1e-10 relative residual)",
        "value": m["steps"] / m["wall"], "unit": "solves/s", "n_gpus": 1,
        "steps": m["steps"], "warmup": args.warmup,
        "ms_per_step": m["wall"] / m["steps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 V-cycle / f64 refinement",
        "data": f"synthetic (16 Gaussian blobs, SplitMix64 seed {seed})",
        "config": {"workload": "poisson-gmg 256^3 on [0,1]^3, 7-point, FixedFace 0, "
                               "omega 2/3 Jacobi 2+2, V-cycle to 1e-10 relative",
                   "levels": 9, "inner_cycles_per_fp64_step": args.inner_cycles,
                   "fp64_steps": m["iters"], "final_rel_residual": m["rel"]},
        "fp64": {"solves_per_s": res["fp64"]["steps"] / res["fp64"]["wall"],
                 "iterations": res["fp64"]["iters"],
                 "final_rel_residual": res["fp64"]["rel"]},
        "speedup_vs_fp64": (m["steps"] / m["wall"]) /
                           (res["fp64"]["steps"] / res["fp64"]["wall"]),
        "mixed_one_cycle_per_step": ({"solves_per_s": res["mixed_inner1"]["steps"] /
                                      res["mixed_inner1"]["wall"],
                                      "fp64_steps": res["mixed_inner1"]["iters"]}
                                     if "mixed_inner1" in res else None),
        "mixed_cell_sweeps": ({"solves_per_s": res["mixed_cell_sweeps"]["steps"] /
                                               res["mixed_cell_sweeps"]["wall"],
                               "same_bits": bool(np.array_equal(sols["mixed_cell_sweeps"],
                                                                sols["mixed"]))}
                              if "mixed_cell_sweeps" in res else None),
        "fp64_kernels": {k: {"launches_per_solve": v[0], "avg_us": v[1] / v[0] * 1e3}
                         for k, v in sorted(res["fp64"]["prof"].items())},
        "error_vs_fp64_solver": float(np.sqrt(np.mean((sols["mixed"] - sols["fp64"]) ** 2)) /
                                      np.sqrt(np.mean(sols["fp64"] ** 2))),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("gmg", dom),
                     "algorithmic_bytes_per_launch": kb[dom], "avg_launch_us": avg_s * 1e6},
        "iteration_roofline": {"algorithmic_bytes_per_iteration": it_bytes,
                               "ms_per_iteration": per_it * 1e3,
                               "achieved_GBps": it_bytes / per_it / 1e9,
                               "frac": it_bytes / per_it / 1e9 / HBM_PEAK_GBS},
        "kernels": {k: {"launches_per_solve": v[0], "avg_us": v[1] / v[0] * 1e3}
                    for k, v in sorted(prof.items())},
    }
    if not args.no_cpu_baseline:
        threads, host = host_cores()
        info, u_cpu = gmg_cpu_baseline(N, seed, threads)
        out["error_vs_oracle"] = float(np.sqrt(np.mean((sols["mixed"] - u_cpu) ** 2)) /
                                       np.sqrt(np.mean(u_cpu ** 2)))
        out["cpu_baseline"] = {"value": 1.0 / info["wall"], "unit": "solves/s",
                               "cores": threads, "kind": "port", "host": host,
                               "sample": f"1 FP64 solve of the same problem by the oracle "
                                         f"({info['iters']} V-cycles, {info['wall']:.1f} s)"}
    if quick:
        return out
    print(json.dumps(out))


def gmg_refined_bench(quick=True):
    """MGSolver<u, rhs, PoissonEquation> on statically refined meshes (the block-mesh
    hierarchy, FP64, SRJ2): the reference's poisson_gmg regression deck (2-D, 64^2 cells in
    32^2 blocks on [-1, 1]^2, [-1, -0.75]^2 refined to logical level 3, use_exact_rhs, 15
    V-cycles to 1e-12 RMS residual) and a 3-D mesh (128^3 cells in 16^3 blocks on [-1, 1]^3,
    the blocks meeting the sphere r = 0.25 refined one level, a seeded Gaussian-blob rhs, to
    1e-10 RMS residual). One step = one solve from u = 0, setup excluded."""
    import numpy as np
    import parthenon_amd as pa
    from parthenon_amd.gmg import MeshMGParams, MeshMGSolver, PoissonEquation

    def build(nx, bnx, lo, hi, boxes=(), spheres=()):
        nd = sum(1 for v in nx if v > 1)
        bcs = tuple(("fixed_face" if d < nd else "periodic") for d in range(3) for _ in range(2))
        m = pa.Mesh(nx, bnx, 2, xmin=lo, xmax=hi, periodic=tuple(0 if d < nd else 1
                                                                 for d in range(3)),
                    bcs=bcs, refine_boxes=boxes, refine_spheres=spheres)
        m.add_field("u", 1, pa.Metadata.FillGhost | pa.Metadata.Independent)
        m.register_refinement_ops("u", "linear")
        m.add_field("rhs", 1)
        m.finalize()
        return m

    def bricks(m, f):  # cell-centre function -> the field's bricks (ghosts too)
        out = np.zeros(m.field_shape("u"))
        n, g = m.n, m.ng
        for gid in range(m.nblocks):
            b = m.block(gid)
            ax = [b["xmin"][d] + (np.arange(n[d] + 2 * g[d]) - g[d] + 0.5) *
                  (b["xmax"][d] - b["xmin"][d]) / n[d] for d in range(3)]
            z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
            out[gid, 0] = f(x, y, z)
        return out

    def run(m, rhs_fn, exact_fn, tol, maxit, steps):
        md = m.mesh_data("base")
        if exact_fn is not None:  # use_exact_rhs: rhs <- A exact (poisson_driver.cpp:87-95)
            md.upload("u", bricks(m, exact_fn))
            pa.exchange_fused(md)
            PoissonEquation().Ax(md, "u", "rhs")
        else:
            md.upload("rhs", bricks(m, rhs_fn))
        zero = np.zeros(m.field_shape("u"))
        s = MeshMGSolver(m, "u", "rhs", MeshMGParams(max_iters=maxit, residual_tolerance=tol,
                                                     precision="fp64", smoother="SRJ2"))
        walls = []
        for q in range(steps + 1):
            md.upload("u", zero)
            s.setup(md)
            m.synchronize()
            t0 = time.perf_counter()
            its, res = s.solve(md)
            m.synchronize()
            walls.append(time.perf_counter() - t0)
        pa.profile_enable(True)
        pa.profile_report()
        md.upload("u", zero)
        s.setup(md)
        s.solve(md)
        prof = pa.profile_report()
        pa.profile_enable(False)
        out = {"blocks": m.nblocks, "levels": s.levels(), "v_cycles": its,
               "rms_residual": res, "ms_per_solve": 1e3 * float(np.median(walls[1:])),
               "solves_per_s": 1.0 / float(np.median(walls[1:])),
               "kernels": {k: {"launches_per_solve": v[0], "avg_us": v[1] / v[0] * 1e3}
                           for k, v in sorted(prof.items())}}
        if exact_fn is not None:
            u = md.download("u")
            ex = bricks(m, exact_fn)
            n, g = m.n, m.ng
            sl = (slice(None), 0, slice(g[2], g[2] + n[2]), slice(g[1], g[1] + n[1]),
                  slice(g[0], g[0] + n[0]))
            out["rms_error"] = float(np.sqrt(np.mean((u[sl] - ex[sl]) ** 2)))
        return out

    deck = build((64, 64, 1), (32, 32, 1), (-1.0, -1.0, 0.0), (1.0, 1.0, 1.0),
                 boxes=[((-1.0, -0.75, -1.0, -0.75, 0.0, 1.0), 3)])
    res_deck = run(deck, None, lambda x, y, z: -np.exp(-10.0 * (x * x + y * y)), 1e-12, 15, 5)
    rng = np.random.default_rng(3)
    blobs = [(rng.uniform(-0.8, 0.8, 3), rng.uniform(0.05, 0.2), rng.choice([-1.0, 1.0]))
             for _ in range(16)]

    def blob_rhs(x, y, z):
        v = np.zeros_like(x)
        for c, w, a in blobs:
            v += a * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (w * w))
        return v
    m3 = build((128, 128, 128), (16, 16, 16), (-1.0,) * 3, (1.0,) * 3,
               spheres=[((0.0, 0.0, 0.0), 0.25, 1)])
    res3 = run(m3, blob_rhs, None, 1e-10, 40, 3)
    return {"metric": "refined-mesh Poisson MG solves/s (FP64, SRJ2)",
            "value": res3["solves_per_s"], "unit": "solves/s", "higher_is_better": True,
            "dtype": "f64", "data": "synthetic (16 Gaussian blobs, numpy seed 3)",
            "config": {"workload": "poisson-gmg 128^3 on [-1,1]^3 in 16^3 blocks, sphere "
                                   "r=0.25 refined one level, FixedFace 0, SRJ2, 1e-10 RMS"},
            "mesh_3d": res3, "regression_deck": res_deck}


def gmg_serve_bench(args, nreq=None, quick=False):
    """Served Poisson solves: seeded Poisson-process arrivals at 50 requests/s, grid sizes
    {32, 64, 128, 192}^3 with weights {0.4, 0.3, 0.2, 0.1}, block size 16^3 or 32^3 (50/50,
    recorded; the solver does not depend on it), blob RHS seeded per request, mixed-precision
    GMG to 1e-10 relative residual. One step = one request; latency = completion - arrival."""
    import numpy as np
    from parthenon_amd.gmg import Server

    nreq = nreq or args.steps
    rng = np.random.default_rng(4)
    arrivals = np.cumsum(rng.exponential(1 / 50.0, nreq))
    sizes = rng.choice([32, 64, 128, 192], size=nreq, p=[0.4, 0.3, 0.2, 0.1])
    blocks = rng.choice([16, 32], size=nreq)
    srv = Server("mixed", max_active=64)
    # warm the memory pool with the request sizes outside the timed stream
    warm = Server("mixed", max_active=8)
    for n in (32, 64, 128, 192):
        warm.submit(int(n), 1)
    warm.run(realtime=False)
    del warm
    ids = [srv.submit(int(n), 1000 + q, float(t)) for q, (n, t) in enumerate(zip(sizes, arrivals))]
    t0 = time.perf_counter()
    srv.run(realtime=True)
    wall = time.perf_counter() - t0
    lat = np.array([srv.result(r)["latency"] for r in ids])
    rel = np.array([srv.result(r)["rel_residual"] for r in ids])
    st = srv.stats()
    # capacity: the same requests all present at t = 0
    cap = Server("mixed", max_active=64)
    for q, n in enumerate(sizes[:min(nreq, 400)]):
        cap.submit(int(n), 1000 + q, 0.0)
    tc = time.perf_counter()
    cap.run(realtime=False)
    cap_wall = time.perf_counter() - tc
    out = {
        "metric": "served Poisson solves/s and latency (mixed-precision GMG, 1e-10)",
        "value": nreq / wall, "unit": "solves/s", "n_gpus": 1, "steps": nreq, "warmup": 0,
        "ms_per_step": wall / nreq * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 V-cycle / f64 refinement",
        "data": "synthetic (Poisson arrivals 50/s, seed 4; blob RHS per request)",
        "config": {"workload": "gmg-serve: sizes {32,64,128,192}^3 p={.4,.3,.2,.1}, "
                               "blocks 16^3/32^3, 50 req/s",
                   "requests": nreq, "block_sizes_drawn": {"16": int(np.sum(blocks == 16)),
                                                           "32": int(np.sum(blocks == 32))}},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3),
                       "max": float(lat.max() * 1e3)},
        "max_rel_residual": float(rel.max()),
        "server": st,
        "capacity_solves_per_s": min(nreq, 400) / cap_wall,
    }
    if quick:
        return out
    print(json.dumps(out))


def remesh_copy_elements(old, new, n, g, comps):
    """Elements the remesh box-copy launch moves: a moved block's whole brick, a refined
    block's coarse-shaped octant of its old parent, a derefined block's 8 restricted octants;
    comps lists the component counts of every moved (container, field) array."""
    E, C, I = (n + 2 * g) ** 3, (n // 2 + 2 * g) ** 3, n ** 3
    olds = set(old)
    per = 0
    for lv, l in new:
        if (lv, l) in olds:
            per += E
        elif lv > 0 and (lv - 1, tuple(x // 2 for x in l)) in olds:
            per += C
        else:
            per += I
    return per * sum(comps)


def adaptive_bench(args, quiet=False):
    """Adaptive Burgers (EvolutionDriver::Execute with refinement = adaptive): the VIBE input's
    criterion (derivative_order_1 on U component 3, refine 0.5 / derefine 0.2, derefine_count
    10) on a 64^3 mesh of 16^3 blocks with numlevel 3. The first cycle's remesh refines every
    block (64 -> 512); its parts are timed separately, then adaptive cycles (cycle + tag +
    remesh check) are compared with plain cycles on the refined mesh."""
    import parthenon_amd as pa
    from parthenon_amd.burgers import DeviceBurgersDriver, make_burgers_mesh
    nx, block, g, ns = (64, 64, 64), (16, 16, 16), 4, 4
    crit = [dict(field="U", comp=3, method="derivative_order_1", refine_tol=0.5,
                 derefine_tol=0.2)]
    mesh = make_burgers_mesh(nx, block, g, ns,
                             adaptive=dict(numlevel=3, derefine_count=10, criteria=crit))
    drv = DeviceBurgersDriver(mesh, cfl=0.4, use_graph=True)
    drv.initialize("sinusoid", seed=2)
    base = mesh.mesh_data("base")
    drv.cycles(1)
    drv.synchronize()
    old = mesh.locs()
    pa.profile_enable(True)
    pa.profile_report()
    t0 = time.perf_counter()
    pa.Refinement.Tag(base)
    t1 = time.perf_counter()
    changed = mesh.adapt()
    t2 = time.perf_counter()
    pa.lib().pth_burgers_after_remesh(mesh._h)
    t3 = time.perf_counter()
    prof = pa.profile_report()
    pa.profile_enable(False)
    new = mesh.locs()
    V = 3 + ns
    copied = remesh_copy_elements(old, new, block[0], g, [V, V, 1])
    kern = {k: {"launches": c, "avg_us": ms / c * 1e3} for k, (c, ms) in sorted(prof.items())}
    cp = prof.get("remesh_copy")
    copy_gbps = 16 * copied / (cp[1] / cp[0] * 1e-3) / 1e9 if cp else None
    # steady state on the refined mesh: plain cycles, then adaptive cycles (tag + remesh check)
    k = max(4, min(args.steps, 40))
    drv.cycles(2)
    drv.synchronize()
    ta = time.perf_counter()
    drv.cycles(k)
    drv.synchronize()
    plain = time.perf_counter() - ta
    ta = time.perf_counter()
    remeshes = drv.adaptive_cycles(k)
    drv.synchronize()
    adaptive = time.perf_counter() - ta
    # the tag's own cost: its launch timed with HIP events, and the synchronous call (launch,
    # flag copy to the host, SetRefinement per block) on the wall clock
    pa.profile_enable(True)
    pa.profile_report()
    tt = time.perf_counter()
    for _ in range(k):
        pa.Refinement.Tag(base)
    tag_wall = (time.perf_counter() - tt) / k
    tprof = pa.profile_report().get("refinement_tag")
    pa.profile_enable(False)
    nb = mesh.nblocks
    zones = nb * block[0] ** 3
    out = {
        "workload": "burgers-adaptive 64^3 mesh of 16^3 blocks, numlevel 3, derivative_order_1 "
                    "on U[3] (refine 0.5, derefine 0.2, derefine_count 10), nghost 4, 3 "
                    "velocities + 4 scalars, CFL 0.4, periodic",
        "first_remesh": {"changed": bool(changed), "blocks_before": len(old),
                         "blocks_after": len(new), "tag_ms": (t1 - t0) * 1e3,
                         "adapt_ms": (t2 - t1) * 1e3, "after_remesh_ms": (t3 - t2) * 1e3,
                         "copy_elements": copied, "remesh_copy_GBps": copy_gbps,
                         "kernels": kern},
        "steady": {"cycles": k, "blocks": nb, "remeshes": remeshes,
                   "adaptive_zone_cycles_per_s": zones * k / adaptive,
                   "plain_zone_cycles_per_s": zones * k / plain,
                   "tag_kernel_us": tprof[1] / tprof[0] * 1e3 if tprof else None,
                   "tag_call_ms": tag_wall * 1e3},
    }
    if not quiet:
        print(json.dumps(out))
    return out


def exchange_bench(args, quiet=False, pack=None, steps=None, warmup=None):
    """Ghost-exchange microbenchmark: 128^3 mesh in 512 blocks of 16^3, nghost 3, 8 FP64
    one-component fields, U(-1,1) interiors; one step = SendBoundBufs + ReceiveBoundBufs +
    SetBounds over all 26 neighbour buffers of every block. Returns the line's dict."""
    import parthenon_amd as pa
    from parthenon_amd.inputs import interior_uniform
    args = argparse.Namespace(**vars(args))
    if pack is not None:
        args.pack = pack
    if steps is not None:
        args.steps = steps
    if warmup is not None:
        args.warmup = warmup

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
    out = {
        "metric": "boundary exchanges/s (Send+Recv+Set, 26 buffers per block)",
        "value": args.steps / wall, "unit": "exchanges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic U(-1,1)",
        "config": {"workload": "exchange 128^3 mesh, 512 blocks of 16^3, nghost 3, 8 fields",
                   "pack": args.pack, "buffers": nbuf, "elements": nel},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": 16 * nel, "avg_launch_us": us[dom]},
        # the buffered figure (pack R+W, unpack R+W) per exchange, also for the fused form
        "exchange_roofline": {"algorithmic_bytes": alg,
                              "achieved_GBps": alg / (wall / args.steps) / 1e9,
                              "frac": alg / (wall / args.steps) / 1e9 / HBM_PEAK_GBS},
        "kernels_avg_us": us}
    del md, m
    if not quiet:
        print(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload",
                    choices=["burgers-amr", "burgers-uniform", "burgers-adaptive", "exchange",
                             "copy", "gmg", "gmg-serve", "gmg-refined"],
                    default="burgers-amr")
    ap.add_argument("--pack", choices=["persistent", "per_buffer", "fused"], default="persistent")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--driver", choices=["graph", "eager", "tasks"], default="graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="burgers: skip timing the plain one-launch-per-buffer task form")
    ap.add_argument("--form", choices=["auto", "sweep", "overlap", "serial"], default="auto",
                    help="burgers device driver stage form: one block-sweep launch per stage; "
                         "plane reconstruction with the exchange on a second stream; plane "
                         "reconstruction, update, exchange on one stream; auto = timed at "
                         "initialization")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="burgers-amr: skip the adaptive remesh measurement in the same line")
    ap.add_argument("--no-exchange", action="store_true",
                    help="leave out the exchange sub-measurement of the default line")
    ap.add_argument("--no-gmg", action="store_true",
                    help="burgers-amr: skip the GMG solve and served-stream sub-objects")
    ap.add_argument("--cpu-cycles", type=int, default=None)
    ap.add_argument("--inner-cycles", type=int, default=2,
                    help="gmg: FP32 V-cycles per FP64 refinement step")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "exchange":
        return exchange_bench(args)
    if args.workload == "copy":
        return copy_bench(args)
    if args.workload == "burgers-adaptive":
        return adaptive_bench(args)
    if args.workload == "gmg":
        if "--steps" not in sys.argv:
            args.steps = 10
        if "--warmup" not in sys.argv:
            args.warmup = 1
        return gmg_bench(args)
    if args.workload == "gmg-refined":
        print(json.dumps(gmg_refined_bench()), flush=True)
        return
    if args.workload == "gmg-serve":
        if "--steps" not in sys.argv:
            args.steps = 2000
        return gmg_serve_bench(args)

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
        form = {} if args.form == "auto" else dict(overlap=args.form == "overlap",
                                                      sweep=args.form == "sweep")
        drv = DeviceBurgersDriver(mesh, cfl=0.4, use_graph=(args.driver == "graph"), **form)
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
    # the block sweep covers every block once per RK stage, in one launch or, split on a
    # refined mesh, in two (the blocks without cross-level channels, then the others): its
    # algorithmic bytes and instruction count are per stage, over the stage's sweep time
    lps = {k: (c / (2 * prof_steps) if k == "block_sweep" else 1.0)
           for k, (c, ms) in prof.items()}
    stage_s = avg_s * lps[dom]
    achieved = kb[dom] / stage_s / 1e9 if dom in kb else None
    form = drv.form() if args.driver != "tasks" else "tasks"
    push_bytes = None
    if form == "sweep":
        # the sweep writes the same-level ghost cells itself; the exchange moves the rest only
        kb.pop("exchange_direct", None)
        push_bytes = 8 * V * same_level_ghost_cells(mesh)
    kernels = {k: {"launches_per_step": c / prof_steps, "avg_us": ms / c * 1e3,
                   "GBps": (kb[k] / (ms / c * 1e-3 * lps[k]) / 1e9) if k in kb else None}
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
                   "block": list(block), "nghost": g, "ncomp": V,
                   "stage_form": form,
                   "tuning_planes_per_wg_and_sweep_waves":
                       drv.tuning() if args.driver != "tasks" else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic("burgers-amr" if amr else "burgers-uniform",
                                                 dom),
                     "algorithmic_bytes_per_stage": kb.get(dom),
                     "launches_per_stage": lps[dom],
                     "avg_launch_us": avg_s * 1e6,
                     "stage_us": stage_s * 1e6,
                     # the block sweep also writes every same-level ghost cell of its output
                     # (the exchange's copies, done in the same launch); counted here apart
                     "push_bytes_per_stage": push_bytes if dom == "block_sweep" else None,
                     "frac_with_push": ((kb[dom] + push_bytes) / stage_s / 1e9 / HBM_PEAK_GBS
                                        if dom == "block_sweep" and push_bytes else None)},
        # the block sweep is bound by FP64 instruction issue, not by HBM: its share of the
        # VALU issue capacity over the live launch time, from the committed instruction count
        "issue_roofline": ({"bound": "valu",
                            "valu_insts_per_stage": measured_valu(
                                "burgers-amr" if amr else "burgers-uniform", dom),
                            "frac": measured_valu("burgers-amr" if amr else "burgers-uniform",
                                                  dom) / (VALU_ISSUE_PER_S * stage_s)}
                           if measured_valu("burgers-amr" if amr else "burgers-uniform", dom)
                           else None),
        # an effective rate, not achieved bandwidth: the bytes the reference's unfused task
        # decomposition would move per cycle over this build's cycle time (the fused kernels
        # skip most of that traffic, so this can exceed what the hardware moves)
        "step_effective_bandwidth": {
            "definition": "reference task-decomposition bytes per cycle / measured cycle time",
            "reference_decomposition_bytes_per_step": sb,
            "effective_GBps": sb / (wall / args.steps) / 1e9,
            "ratio_to_hbm_peak": sb / (wall / args.steps) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
    }
    if args.driver != "tasks" and not args.no_plain:
        # the plain HIP form of the same step: the reference's task order call by call, one
        # kernel launch per boundary buffer, no graph, no fused kernels
        drv = None
        pmesh = make_burgers_mesh(nx, block, g, ns, refine_spheres=spheres)
        pmesh.set_pack_mode(pa.PackMode.per_buffer)
        pd = BurgersDriver(pmesh, cfl=0.4)
        pd.initialize("sinusoid", seed=seed)
        pd.cycle()
        pmesh.synchronize()
        pc = 3
        tp = time.perf_counter()
        for _ in range(pc):
            pd.cycle()
        pmesh.synchronize()
        pwall = time.perf_counter() - tp
        pv = nb * n ** 3 * pc / pwall
        out["plain_form"] = {
            "value": pv, "unit": "zone-cycles/s", "ms_per_step": pwall / pc * 1e3,
            "speedup": value / pv,
            "form": "task-by-task calls in the reference order, one launch per boundary "
                    f"buffer, eager ({pc} cycles after 1 warm-up)"}
        del pd, pmesh
    if amr and args.driver != "tasks" and not args.no_adaptive:
        out["adaptive"] = adaptive_bench(args, quiet=True)
    if amr and args.driver != "tasks" and not args.no_gmg:
        # the metric's second half on the same GPU in the same run: the 256^3 mixed GMG solve
        # (with its FP64 twin and the oracle's FP64 solve) and the served stream of 2000 requests
        drv = None
        out["gmg"] = gmg_bench(argparse.Namespace(steps=5, warmup=1, inner_cycles=2,
                                                  no_cpu_baseline=args.no_cpu_baseline),
                               quick=True)
        out["gmg_serve"] = gmg_serve_bench(args, nreq=2000, quick=True)  # cfg 4: 2000 requests
        out["gmg_refined"] = gmg_refined_bench()
    if amr and args.driver != "tasks" and not args.no_exchange:
        # the ghost-exchange microbenchmark (128^3 in 512 blocks of 16^3, nghost 3, 8 fields)
        # in its fused (direct copy) and buffered (pack + unpack) forms
        drv = None
        out["exchange"] = {
            form: exchange_bench(args, quiet=True, pack=pk, steps=100, warmup=5)
            for form, pk in (("fused", "fused"), ("buffered", "persistent"))}
    if not args.no_cpu_baseline:
        threads, host = host_cores()
        cw = cpu_baseline(nx, block, g, ns, spheres, seed, cpu_cycles, threads)
        out["cpu_baseline"] = {
            "value": nb * n ** 3 * cpu_cycles / cw, "unit": "zone-cycles/s",
            "cores": threads, "kind": "port", "host": host,
            "sample": f"median of 3 timings of {cpu_cycles} consecutive RK2 cycles of the same "
                      f"workload after 1 warm-up cycle, oracle C++ restatement with OpenMP over "
                      f"meshblocks ({cw:.2f} s per timing)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
