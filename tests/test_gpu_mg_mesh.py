"""GPU parity of the multigrid path over MeshData through the C ABI.

* Solver field utilities (src/solvers/solver_utils.hpp:151-301): copies, weighted sums and
  zeroing are bit-exact against numpy's IEEE evaluation of the same operations; the dot
  product is within 1e-14 relative of the correctly rounded sum (math.fsum) and gives the
  same bits on every call.
* PoissonEquation::Ax on bricks (example/poisson_gmg/poisson_equation.hpp:41-205): bit-equal
  to the oracle's 7-point operator on the unit cube for D = 1, alpha = 0; for other D and alpha
  through linearity (1e-13 relative).
* MGSolver<u, rhs, PoissonEquation> over MeshData (src/solvers/mg_solver.hpp:73-208): in FP64
  on the unit cube its solution is bit-identical to the level-array solver's after the same
  number of V-cycles and within 1e-12 relative RMS of the oracle's; in mixed precision on a
  scaled domain its RMS residual, recomputed from the solution with the brick operator, is
  below the tolerance and the solution is within 1e-7 relative RMS of the oracle's FP64
  solution of the same problem.
* The C++ driver shaped like example/poisson_gmg/poisson_driver.cpp:64-133 (tests/cpp/
  poisson_driver.cpp): rhs = A exact bit-equal to the oracle's operator, the solution against
  the oracle, and the reference's own pass thresholds (RMS residual <= 1e-10, RMS error
  <= 1e-12, example/poisson_gmg/main.cpp:52) on the deck's [-1, 1] domain."""
import math
import os
import subprocess

import numpy as np
import pytest

import parthenon_amd as pa
from oracle_lib import OracleGmg
from parthenon_amd.gmg import (MGParams, MGSolver, MeshMGParams, MeshMGSolver,
                               PoissonEquation, add_fields_and_store, copy_data, dot_product,
                               set_to_zero)

pytestmark = pytest.mark.gpu


def poisson_mesh(N, block, side=1.0, ng=2, extra=(), bcs=("fixed_face",) * 6):
    lo = -0.5 * side
    m = pa.Mesh((N,) * 3, (block,) * 3, ng, xmin=(lo,) * 3, xmax=(-lo,) * 3,
                periodic=(0, 0, 0), bcs=bcs)
    m.add_field("u", 1, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.add_field("rhs", 1)
    for name, nc in extra:
        m.add_field(name, nc)
    m.finalize()
    return m


def to_bricks(m, g, name="u", fill=0.0):
    """Global [k][j][i] array -> the field's bricks (ghosts = fill)."""
    out = np.full(m.field_shape(name), fill)
    n, ng = m.n, m.ng
    for gid in range(m.nblocks):
        lx, ly, lz = m.block(gid)["l"]
        out[gid, 0, ng[2]:ng[2] + n[2], ng[1]:ng[1] + n[1], ng[0]:ng[0] + n[0]] = g[
            lz * n[2]:(lz + 1) * n[2], ly * n[1]:(ly + 1) * n[1], lx * n[0]:(lx + 1) * n[0]]
    return out


def from_bricks(m, f, N):
    g = np.empty((N, N, N))
    n, ng = m.n, m.ng
    for gid in range(m.nblocks):
        lx, ly, lz = m.block(gid)["l"]
        g[lz * n[2]:(lz + 1) * n[2], ly * n[1]:(ly + 1) * n[1], lx * n[0]:(lx + 1) * n[0]] = f[
            gid, 0, ng[2]:ng[2] + n[2], ng[1]:ng[1] + n[1], ng[0]:ng[0] + n[0]]
    return g


def rel_rms(a, b):
    return np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))


def test_solver_utils_match_numpy():
    print("case solver utils", flush=True)
    m = poisson_mesh(32, 16, extra=(("a", 2), ("b", 2), ("c", 2)))
    md = m.mesh_data("base")
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, m.field_shape("a"))
    b = rng.uniform(-1, 1, m.field_shape("b"))
    c = rng.uniform(-1, 1, m.field_shape("c"))
    for name, v in (("a", a), ("b", b), ("c", c)):
        md.upload(name, v)
    copy_data(md, "a", "c")
    assert np.array_equal(md.download("c").view(np.uint64), a.view(np.uint64))
    add_fields_and_store(md, "a", "b", "c", 0.5, -2.0)
    assert np.array_equal(md.download("c"), 0.5 * a + (-2.0) * b)
    add_fields_and_store(md, "c", "b", "c", 1.0, 1.0)  # out aliasing an input
    assert np.array_equal(md.download("c"), (0.5 * a + (-2.0) * b) + b)
    set_to_zero(md, "a")
    got = md.download("a")
    ng, n = m.ng, m.n
    inner = (slice(None), slice(None), slice(ng[2], ng[2] + n[2]), slice(ng[1], ng[1] + n[1]),
             slice(ng[0], ng[0] + n[0]))
    assert not got[inner].any()
    keep = np.ones(got.shape, dtype=bool)
    keep[inner] = False
    assert np.array_equal(got[keep], a[keep])
    d1, d2 = dot_product(md, "b", "c"), dot_product(md, "b", "c")
    cc = md.download("c")
    exact = math.fsum((b[inner] * cc[inner]).ravel())
    assert abs(d1 - exact) <= 1e-14 * abs(exact)
    assert np.float64(d1).view(np.uint64) == np.float64(d2).view(np.uint64)
    with pytest.raises(RuntimeError):
        copy_data(md, "a", "u")  # component counts differ


def test_poisson_ax_matches_oracle_operator():
    print("case poisson ax", flush=True)
    N = 32
    m = poisson_mesh(N, 16, extra=(("ax", 1), ("ax2", 1)))
    md = m.mesh_data("base")
    x = np.random.default_rng(3).uniform(-1, 1, (N, N, N))
    md.upload("u", to_bricks(m, x, fill=np.nan))
    pa.exchange_fused(md)  # neighbours and the FixedFace conditions fill u's ghosts
    PoissonEquation().Ax(md, "u", "ax")
    got = from_bricks(m, md.download("ax"), N)
    assert np.array_equal(got, OracleGmg(N).apply(x))
    # D and alpha: A_{D,alpha} x = D A_1 x - alpha x
    PoissonEquation(D=2.5, diagonal_alpha=0.3).Ax(md, "u", "ax2")
    g2 = from_bricks(m, md.download("ax2"), N)
    ref = 2.5 * got - 0.3 * x
    assert np.max(np.abs(g2 - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("N,block", [(32, 16), (48, 16)])
def test_mesh_solver_fp64_matches_level_solver_and_oracle(N, block):
    print(f"case mesh mg fp64 N={N} block={block}", flush=True)
    g = OracleGmg(N)
    b = g.rhs_blobs(11)
    m = poisson_mesh(N, block)
    md = m.mesh_data("base")
    md.upload("rhs", to_bricks(m, b, "rhs"))
    md.upload("u", to_bricks(m, np.zeros((N, N, N))))
    tol = 1e-10 * np.linalg.norm(b) / math.sqrt(N ** 3)  # RMS tolerance = 1e-10 relative
    s = MeshMGSolver(m, "u", "rhs", MeshMGParams(precision="fp64", residual_tolerance=tol,
                                                  max_iters=60))
    s.setup(md)
    it, rms = s.solve(md)
    assert rms < tol and it > 1, (it, rms, tol)
    u = from_bricks(m, md.download("u"), N)
    # the level-array solver on the same right-hand side, the same number of V-cycles
    ref = MGSolver(N, "fp64", MGParams(max_iters=it, residual_tolerance=0.0))
    ref.set_rhs(b)
    its, _ = ref.solve()
    assert its == it
    assert np.array_equal(u, ref.solution())
    hist = s.history()
    rel = ref.history()
    assert np.allclose(hist, rel * np.linalg.norm(b) / math.sqrt(N ** 3), rtol=1e-13, atol=0)
    # the oracle, the same number of cycles
    g.set_rhs(b)
    g.solve(tol=0.0, maxit=it)
    assert rel_rms(u, g.u()) < 1e-12
    # u's ghosts were refilled: the brick operator's residual agrees with the solver's
    m2 = md.download("u")
    assert not np.isnan(m2).any()


def test_mesh_solver_mixed_on_scaled_domain():
    print("case mesh mg mixed side=2 D=2.5", flush=True)
    N, side, D = 64, 2.0, 2.5
    m = poisson_mesh(N, 16, side=side, extra=(("ax", 1),))
    md = m.mesh_data("base")
    h = side / N
    c = -0.5 * side + (np.arange(N) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    rhs = np.where(np.sqrt(x * x + y * y + z * z) < 0.5, 1.0, 0.0)
    md.upload("rhs", to_bricks(m, rhs, "rhs"))
    md.upload("u", to_bricks(m, np.zeros((N, N, N))))
    eq = PoissonEquation(D=D)
    tol = 1e-10
    s = MeshMGSolver(m, "u", "rhs", MeshMGParams(residual_tolerance=tol, max_iters=60), eq)
    s.setup(md)
    it, rms = s.solve(md)
    assert rms < tol, (it, rms)
    # independent residual: rhs - A u with the brick operator on the refilled ghosts
    eq.Ax(md, "u", "ax")
    r = rhs - from_bricks(m, md.download("ax"), N)
    rms2 = math.sqrt(np.mean(r * r))
    assert rms2 < 1.05 * tol and abs(rms2 - rms) <= 1e-2 * rms, (rms2, rms)
    # against the oracle's FP64 solution of the same problem on the unit cube
    g = OracleGmg(N)
    g.set_rhs(rhs * side * side / D)
    g.solve(tol=1e-13, maxit=100)
    u = from_bricks(m, md.download("u"), N)
    assert rel_rms(u, g.u()) < 1e-7
    # a second solve from the converged solution: one V-cycle (the reference's sublist runs at
    # least one), still below the tolerance
    s.setup(md)
    it2, rms3 = s.solve(md)
    assert it2 == 1 and rms3 < tol


@pytest.mark.parametrize("what", ["outflow", "alpha", "noncubic"])
def test_mesh_solver_rejects_unsupported(what):
    print(f"case mesh mg rejects {what}", flush=True)
    if what == "outflow":
        m = poisson_mesh(32, 16, bcs=("fixed_face",) * 5 + ("outflow",))
    elif what == "noncubic":
        m = pa.Mesh((32, 32, 16), (16, 16, 16), 2, periodic=(0, 0, 0),
                    bcs=("fixed_face",) * 6)
        m.add_field("u", 1, pa.Metadata.FillGhost)
        m.add_field("rhs", 1)
        m.finalize()
    else:
        m = poisson_mesh(32, 16)
    eq = PoissonEquation(diagonal_alpha=0.5 if what == "alpha" else 0.0)
    with pytest.raises(RuntimeError):
        MeshMGSolver(m, "u", "rhs", MeshMGParams(), eq)


def _driver(args, tmp_path):
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "poisson_driver")
    out = str(tmp_path / "poisson.bin")
    r = subprocess.run([exe] + [str(a) for a in args[:2]] + [out] + [str(a) for a in args[2:]],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    N = int(args[0])
    raw = np.fromfile(out, dtype=np.float64)
    res, err, its = raw[:3]
    u, rhs, ex = (raw[3 + q * N ** 3:3 + (q + 1) * N ** 3].reshape(N, N, N) for q in range(3))
    return res, err, int(its), u, rhs, ex


def test_cpp_poisson_driver_matches_oracle(tmp_path):
    print("case cpp poisson driver unit cube fp64", flush=True)
    N = 32
    res, err, its, u, rhs, ex = _driver([N, 16, "fp64", "exact", 1.0, 1.0, 1e-11], tmp_path)
    g = OracleGmg(N)
    # rhs = A exact, with the exchange's FixedFace ghosts: the oracle's operator, same bits
    assert np.array_equal(rhs, g.apply(ex))
    g.set_rhs(rhs)
    g.solve(tol=0.0, maxit=its)
    assert rel_rms(u, g.u()) < 1e-12
    assert res < 1e-11 and its >= 1
    assert abs(err - math.sqrt(np.mean((ex - u) ** 2))) <= 1e-6 * err


def test_cpp_poisson_driver_reference_thresholds(tmp_path):
    # the regression deck's domain [-1, 1] and the pass test of example/poisson_gmg/main.cpp:52
    print("case cpp poisson driver [-1,1]^3 mixed", flush=True)
    res, err, its, u, rhs, ex = _driver([64, 16, "mixed", "exact", 2.0, 1.0, 1e-12], tmp_path)
    assert res <= 1e-10 and err <= 1e-12, (res, err, its)
