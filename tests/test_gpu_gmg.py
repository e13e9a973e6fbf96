"""GPU parity of the geometric-multigrid Poisson solver through the C ABI.

The FP64 solver runs the oracle's algorithm operation for operation, so its residual history
and solution match the oracle's to round-off of the norm summation (stated: 1e-12 relative
on every history entry and on the solution). The mixed solver (FP32 V-cycles inside FP64
iterative refinement) must reach the same 1e-10 relative residual, measured in FP64, and its
solution must agree with the oracle's FP64 solution within 1e-8 relative RMS (the two
solutions each satisfy the discrete equation to 1e-10; the condition number of the 7-point
operator at these sizes bounds their distance). Served solves must be bit-identical alone and
in a batch."""
import numpy as np
import pytest

from oracle_lib import OracleGmg
from parthenon_amd.gmg import MGParams, MGSolver, Server

pytestmark = pytest.mark.gpu


def oracle_solve(N, seed=None, rhs=None, tol=1e-10):
    g = OracleGmg(N)
    b = g.rhs_blobs(seed) if rhs is None else rhs
    if rhs is not None:
        g.set_rhs(rhs)
    it, hist = g.solve(tol=tol, maxit=100)
    return g, b, it, hist


def rel_rms(a, b):
    return np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))


@pytest.mark.parametrize("N", [24, 32, 48])
def test_fp64_solver_matches_oracle(N):
    print(f"case gmg fp64 N={N}", flush=True)
    g, b, it, hist = oracle_solve(N, seed=3)
    s = MGSolver(N, "fp64")
    assert s.levels == g.levels
    # the device blob RHS agrees with the host one to round-off (the device exp and the host
    # exp may differ in the last bit of each of the 16 terms)
    s.set_rhs_blobs(3)
    db = s.rhs()
    assert np.max(np.abs(db - b)) <= 1e-14 * np.max(np.abs(b))
    # the solver itself is compared on the oracle's RHS
    s.set_rhs(b)
    its, rel = s.solve()
    h = s.history()
    assert its == it, (its, it)
    assert np.max(np.abs(h - hist) / hist) < 1e-12
    assert rel_rms(s.solution(), g.u()) < 1e-12


@pytest.mark.parametrize("N", [24, 32, 64])
def test_mixed_solver_reaches_tolerance_and_matches_oracle(N):
    print(f"case gmg mixed N={N}", flush=True)
    g, b, it, hist = oracle_solve(N, seed=5)
    s = MGSolver(N, "mixed")
    s.set_rhs_blobs(5)
    its, rel = s.solve()
    assert rel < 1e-10 and its <= it + 3, (its, it, rel)
    u = s.solution()
    # independent FP64 check of the residual with the oracle's operator
    r = b - g.apply(u)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
    assert rel_rms(u, g.u()) < 1e-8


def test_mixed_known_answer():
    # rhs = A u* for a known u* (the reference's use_exact_rhs mode)
    N = 32
    g = OracleGmg(N)
    x = (np.arange(N) + 0.5) / N
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    ustar = np.sin(np.pi * xx) * np.sin(2 * np.pi * y) * z * (1 - z)
    b = g.apply(ustar)
    s = MGSolver(N, "mixed")
    s.set_rhs(b)
    its, rel = s.solve()
    assert rel < 1e-10
    assert rel_rms(s.solution(), ustar) < 1e-8


def test_mixed_192_hierarchy_odd_coarsest():
    # 192 = 3 * 2^6: levels 192 ... 3; residual checked with the oracle's FP64 operator
    print("case gmg mixed N=192", flush=True)
    s = MGSolver(192, "mixed")
    assert s.levels == 7
    s.set_rhs_blobs(9)
    its, rel = s.solve()
    assert rel < 1e-10, (its, rel)
    g = OracleGmg(192)
    b = s.rhs()
    r = b - g.apply(s.solution())
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_served_solves_identical_alone_and_batched():
    print("case gmg server batch vs alone", flush=True)
    reqs = [(16, 1), (32, 2), (24, 3), (64, 4), (32, 5), (48, 6)]
    batched = Server("mixed", max_active=8, keep_solutions=True)
    ids = [batched.submit(n, seed) for n, seed in reqs]
    batched.run(realtime=False)
    st = batched.stats()
    assert st["max_batch"] == len(reqs)
    for (n, seed), rid in zip(reqs, ids):
        res = batched.result(rid)
        assert res["rel_residual"] < 1e-10
        alone = MGSolver(n, "mixed")
        alone.set_rhs_blobs(seed)
        its, rel = alone.solve()
        assert its == res["iters"]
        assert np.array_equal(batched.solution(rid), alone.solution()), (n, seed)


def test_served_stream_in_real_time():
    print("case gmg server realtime", flush=True)
    rng = np.random.default_rng(4)
    srv = Server("mixed", max_active=16)
    t = 0.0
    ids = []
    for q in range(24):
        t += rng.exponential(1 / 100.0)
        n = int(rng.choice([16, 32, 64]))
        ids.append(srv.submit(n, 100 + q, t))
    srv.run(realtime=True)
    lat = []
    for rid in ids:
        r = srv.result(rid)
        assert r["rel_residual"] < 1e-10
        assert r["admit"] >= r["arrival"] - 1e-9 and r["done"] >= r["admit"]
        lat.append(r["latency"])
    assert np.all(np.array(lat) > 0)


def test_mixed_256_blob_problem_at_full_size():
    # the 256^3 problem of the GMG benchmark (16 blobs, seed 3): the mixed solver's residual,
    # recomputed in FP64 with the oracle's operator, and its distance to the FP64 oracle solve
    print("case gmg mixed N=256 seed 3", flush=True)
    N = 256
    g, b, it, hist = oracle_solve(N, seed=3)
    s = MGSolver(N, "mixed", MGParams(inner_cycles=2))
    assert s.levels == g.levels == 9
    s.set_rhs(b)
    its, rel = s.solve()
    assert rel < 1e-10, (its, rel)
    u = s.solution()
    r = b - g.apply(u)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
    err = rel_rms(u, g.u())
    print(f"N=256: {its} FP64 steps (oracle {it} V-cycles), error vs oracle {err:.2e}", flush=True)
    assert err < 1e-8


def test_served_stream_with_large_grids_matches_solves_alone():
    # a seeded arrival stream that includes 128^3 and 192^3 requests, served in real time with
    # overlapping solves; every result equals the same request solved alone, bit for bit
    print("case gmg server stream with 128^3 and 192^3", flush=True)
    rng = np.random.default_rng(7)
    sizes = [32, 128, 64, 192, 32, 128, 64, 32, 192, 64, 32, 128]
    t, reqs = 0.0, []
    for q, n in enumerate(sizes):
        t += rng.exponential(1 / 50.0)
        reqs.append((n, 300 + q, t))
    srv = Server("mixed", max_active=16, keep_solutions=True)
    ids = [srv.submit(n, seed, at) for n, seed, at in reqs]
    srv.run(realtime=True)
    st = srv.stats()
    print(f"server stats {st}", flush=True)
    assert st["max_batch"] >= 2
    for (n, seed, _), rid in zip(reqs, ids):
        res = srv.result(rid)
        assert res["rel_residual"] < 1e-10
        alone = MGSolver(n, "mixed")
        alone.set_rhs_blobs(seed)
        its, rel = alone.solve()
        assert its == res["iters"], (n, seed)
        assert np.array_equal(srv.solution(rid), alone.solution()), (n, seed)


def test_admission_by_free_device_memory():
    # with a reserve that leaves room for two 64^3 hierarchies in the device's free memory,
    # at most two of six requests present at t = 0 are active together; results are unchanged
    import gc
    import parthenon_amd as pa
    print("case gmg server admission by free HBM", flush=True)
    gc.collect()
    probe = Server("mixed", max_active=8)
    probe.submit(64, 1)
    probe.run(realtime=False)
    per = probe.stats()["bytes_held"]
    del probe
    srv = Server("mixed", max_active=8, keep_solutions=True)
    free, total = pa.device_memory()
    assert 0 < free <= total
    pa.lib().pth_gmg_server_set_hbm_reserve(srv._h, int(free - 2.5 * per))
    ids = [srv.submit(64, 40 + q, 0.0) for q in range(6)]
    srv.run(realtime=False)
    st = srv.stats()
    assert st["max_batch"] == 2, st
    for q, rid in enumerate(ids):
        alone = MGSolver(64, "mixed")
        alone.set_rhs_blobs(40 + q)
        alone.solve()
        assert np.array_equal(srv.solution(rid), alone.solution())
