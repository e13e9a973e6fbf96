"""GPU parity of MGSolver<u, rhs, PoissonEquation> on statically refined and 2-D meshes (the
block-mesh hierarchy), through the C ABI.

* The reference's own regression deck (tst/regression/test_suites/poisson_gmg/
  parthinput.poisson): 2-D, 64^2 cells on [-1, 1]^2 in 32^2 blocks, static refinement of
  [-1, -0.75]^2 to logical level 3 (three levels above the base grid: a 2-D forest has root
  level 0), FixedFace faces, SRJ2,
  max_iterations 15, residual_tolerance 1e-12, use_exact_rhs (rhs = A exact with
  exact = -exp(-10 r^2), poisson_driver.cpp:87-95, parthenon_app_inputs.cpp:53-75). rhs is
  bit-equal to the oracle's leaf operator applied to exact; the solve passes the reference's
  thresholds (RMS residual <= 1e-10, RMS error <= 1e-12, example/poisson_gmg/main.cpp:52)
  and matches the oracle's BiCGSTAB solution of the same leaf problem.
* The deck's own right-hand side (1 inside r < 0.5, parthenon_app_inputs.cpp:53-68) and a 3-D
  sphere-refined mesh: the solution against the oracle's, 1e-9 relative RMS, residual
  recomputed with the oracle's leaf operator below the tolerance.
* tests/cpp/poisson_deck_driver on the deck file itself (tests/golden/parthinput.poisson):
  the same thresholds and the oracle's solution, through the C++ task-level mirror.
"""
import os
import subprocess

import numpy as np
import pytest

import parthenon_amd as pa
from oracle_lib import OracleMesh
from parthenon_amd.gmg import MeshMGParams, MeshMGSolver, PoissonEquation

pytestmark = pytest.mark.gpu

# <parthenon/static_refinement0> of the deck: logical level 3; a 2-D forest has root level 0
# (forest.cpp:74-80), so three levels above the base grid
DECK_BOX = ((-1.0, -0.75, -1.0, -0.75, 0.0, 1.0), 3)


def make_meshes(nx, bnx, xmin, xmax, boxes=(), spheres=(), ng=2, face_D=False):
    nd = sum(1 for v in nx if v > 1)
    bcs = tuple(("fixed_face" if d < nd else "periodic") for d in range(3) for _ in range(2))
    periodic = tuple(0 if d < nd else 1 for d in range(3))
    m = pa.Mesh(nx, bnx, ng, xmin=xmin, xmax=xmax, periodic=periodic, bcs=bcs,
                refine_boxes=boxes, refine_spheres=spheres)
    m.add_field("u", 1, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.register_refinement_ops("u", "linear")  # poisson_package.cpp:122-126
    m.add_field("rhs", 1)
    m.add_field("Ax", 1)
    if face_D:
        m.add_field("D", 3, 0)  # the face field D: component d on each cell's lower face
    m.finalize()
    om = OracleMesh(nx, bnx, ng, xmin=xmin, xmax=xmax, periodic=periodic, bcs=bcs,
                    boxes=boxes, spheres=spheres)
    assert om.nblocks == m.nblocks
    for gid in range(m.nblocks):
        a, b = m.block(gid), om.block(gid)
        assert (a["level"], tuple(a["l"])) == (b["level"], tuple(b["l"]))
    return m, om


def centres(m):
    """Cell centres of every interior cell, block order then [k][j][i]."""
    n = m.n
    out = []
    for gid in range(m.nblocks):
        b = m.block(gid)
        lo, hi = b["xmin"], b["xmax"]
        ax = [lo[d] + (np.arange(n[d]) + 0.5) * (hi[d] - lo[d]) / n[d] for d in range(3)]
        z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        out.append(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1))
    return np.concatenate(out)


def to_bricks(m, v, name="u", fill=0.0):
    out = np.full(m.field_shape(name), fill)
    n, g = m.n, m.ng
    per = n[0] * n[1] * n[2]
    for gid in range(m.nblocks):
        out[gid, 0, g[2]:g[2] + n[2], g[1]:g[1] + n[1], g[0]:g[0] + n[0]] = \
            v[gid * per:(gid + 1) * per].reshape(n[2], n[1], n[0])
    return out


def from_bricks(m, f):
    n, g = m.n, m.ng
    return np.concatenate([f[gid, 0, g[2]:g[2] + n[2], g[1]:g[1] + n[1], g[0]:g[0] + n[0]]
                           .ravel() for gid in range(m.nblocks)])


def rms(v):
    return float(np.sqrt(np.mean(np.square(v))))


def deck(face_D=False):
    return make_meshes((64, 64, 1), (32, 32, 1), (-1.0, -1.0, 0.0), (1.0, 1.0, 1.0),
                       boxes=[DECK_BOX], face_D=face_D)


def deck_face_D(m, interior_D, exterior_D):
    """The problem generator's D (parthenon_app_inputs.cpp:58-75, 2-D): interior_D on faces
    inside the two rectangles, exterior_D elsewhere, as [block][3] bricks."""
    def inside(x, y):
        in1 = (x < -0.25) & (x > -0.75) & (y < 0.5) & (y > -0.5)
        in2 = (x < 0.5) & (x > -0.75) & (y < -0.25) & (y > -0.75)
        return in1 | in2
    n, g = m.n, m.ng
    e = tuple(n[d] + 2 * g[d] for d in range(3))
    out = np.zeros((m.nblocks, 3, e[2], e[1], e[0]))
    for gid in range(m.nblocks):
        b = m.block(gid)
        dx = [(b["xmax"][d] - b["xmin"][d]) / n[d] for d in range(3)]
        cc = [b["xmin"][d] + (np.arange(e[d]) - g[d] + 0.5) * dx[d] for d in range(3)]
        for d in range(3):
            ax = [a.copy() for a in cc]
            ax[d] = ax[d] - 0.5 * dx[d]
            z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
            out[gid, d] = np.where(inside(x, y), interior_D, exterior_D)
    return out


def test_deck_use_exact_rhs_passes_reference_thresholds():
    print("case poisson_gmg deck, use_exact_rhs", flush=True)
    m, om = deck()
    md = m.mesh_data("base")
    c = centres(m)
    r = np.sqrt(c[:, 0] ** 2 + c[:, 1] ** 2)  # ndim = 2: x3 does not enter the radius
    exact = -np.exp(-10.0 * r * r)
    # rhs <- A exact: CopyData<exact, u>, boundary exchange, Ax<u, rhs>
    md.upload("u", to_bricks(m, exact, fill=np.nan))
    pa.exchange_fused(md)
    PoissonEquation().Ax(md, "u", "rhs")
    rhs = from_bricks(m, md.download("rhs"))
    ref_rhs = om.poisson_apply(exact)
    assert np.array_equal(rhs, ref_rhs), np.max(np.abs(rhs - ref_rhs))
    md.upload("u", np.zeros(m.field_shape("u")))
    p = MeshMGParams(max_iters=15, residual_tolerance=1e-12, precision="fp64",
                     smoother="SRJ2")
    s = MeshMGSolver(m, "u", "rhs", p)
    # the reference's GMG levels -5 .. 3 (mesh-gmg.cpp:103-118): the leaf mesh, the mesh capped
    # at logical levels 2, 1, 0, then blocks of 16^2 .. 1^2 cells
    assert m.nblocks == 13 and s.levels() == 9
    s.setup(md)
    it, res = s.solve(md)
    h = s.history()
    print(f"deck: {it} V-cycles, rms residual {res:.3e}, history {h}", flush=True)
    u = from_bricks(m, md.download("u"))
    err = rms(u - exact)
    print(f"deck: rms error {err:.3e}", flush=True)
    assert res <= 1e-10 and err <= 1e-12  # example/poisson_gmg/main.cpp:52
    assert res == pytest.approx(rms(ref_rhs - om.poisson_apply(u)), rel=1e-6, abs=1e-15)
    xo, _, _ = om.poisson_solve(ref_rhs, tol=1e-13)
    assert rms(u - xo) <= 1e-12
    assert np.all(np.diff(h[:it]) < 0)  # monotone until the tolerance


@pytest.mark.parametrize("smoother", ["SRJ2", "jacobi"])
def test_deck_rhs_matches_oracle(smoother):
    print(f"case poisson_gmg deck rhs, {smoother}", flush=True)
    m, om = deck()
    md = m.mesh_data("base")
    c = centres(m)
    b = np.where(np.sqrt(c[:, 0] ** 2 + c[:, 1] ** 2) < 0.5, 1.0, 0.0)
    md.upload("rhs", to_bricks(m, b, name="rhs"))
    md.upload("u", np.zeros(m.field_shape("u")))
    p = MeshMGParams(max_iters=60, residual_tolerance=1e-11, precision="fp64",
                     smoother=smoother)
    s = MeshMGSolver(m, "u", "rhs", p)
    s.setup(md)
    it, res = s.solve(md)
    print(f"deck rhs {smoother}: {it} V-cycles, rms residual {res:.3e}", flush=True)
    assert res <= 1e-11
    u = from_bricks(m, md.download("u"))
    assert rms(b - om.poisson_apply(u)) <= 1.01e-11
    xo, _, _ = om.poisson_solve(b, tol=1e-13)
    assert rms(u - xo) <= 1e-9 * rms(xo)
    # u's ghosts were refilled by the boundary exchange after the solve
    full = md.download("u")
    md.upload("u", full)
    pa.exchange_fused(md)
    assert np.array_equal(md.download("u"), full)


def test_sphere_refined_3d_matches_oracle():
    print("case 3-D sphere-refined mesh", flush=True)
    m, om = make_meshes((32, 32, 32), (8, 8, 8), (-0.5,) * 3, (0.5,) * 3,
                        spheres=[((0.1, -0.05, 0.0), 0.15, 1)])
    md = m.mesh_data("base")
    c = centres(m)
    rng = np.random.default_rng(11)
    b = np.zeros(len(c))
    for _ in range(4):
        x0 = rng.uniform(-0.3, 0.3, 3)
        b += rng.choice([-1.0, 1.0]) * np.exp(-np.sum((c - x0) ** 2, axis=1) / 0.01)
    md.upload("rhs", to_bricks(m, b, name="rhs"))
    md.upload("u", np.zeros(m.field_shape("u")))
    p = MeshMGParams(max_iters=60, residual_tolerance=1e-10, precision="fp64",
                     smoother="SRJ2")
    s = MeshMGSolver(m, "u", "rhs", p)
    s.setup(md)
    it, res = s.solve(md)
    print(f"3-D: {it} V-cycles, rms residual {res:.3e}, levels {s.levels()}", flush=True)
    assert res <= 1e-10
    u = from_bricks(m, md.download("u"))
    assert rms(b - om.poisson_apply(u)) <= 1.01e-10
    xo, _, _ = om.poisson_solve(b, tol=1e-13)
    assert rms(u - xo) <= 1e-8 * rms(xo)


def test_mixed_precision_is_refused_on_refined_meshes():
    m, _ = deck()
    with pytest.raises(RuntimeError, match="mixed precision"):
        MeshMGSolver(m, "u", "rhs", MeshMGParams(precision="mixed"))


def test_cpp_deck_driver_runs_reference_deck(tmp_path):
    """tests/cpp/poisson_deck_driver reads the reference's regression deck (copied unchanged
    into tests/golden) and runs PoissonDriver's task collection through the C++ mirror."""
    print("case cpp deck driver on parthinput.poisson", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "cpp", "poisson_deck_driver")
    out = str(tmp_path / "deck.bin")
    r = subprocess.run([exe, os.path.join(here, "golden", "parthinput.poisson"), out],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    raw = np.fromfile(out, dtype=np.float64)
    res, err, its, nb, per = raw[:5]
    n = int(nb) * int(per)
    u, rhs, exact = (raw[5 + q * n:5 + (q + 1) * n] for q in range(3))
    # the reference's pass test (example/poisson_gmg/main.cpp:52), within its 15 V-cycles
    assert res <= 1e-10 and err <= 1e-12 and its <= 15
    _, om = deck()
    assert np.array_equal(rhs, om.poisson_apply(exact))
    assert abs(err - rms(u - exact)) <= 1e-6 * err
    xo, _, _ = om.poisson_solve(rhs, tol=1e-13)
    assert rms(u - xo) <= 1e-12


def test_deck_with_face_field_D_matches_oracle():
    """The deck with interior_D = 10, exterior_D = 1 (the package's face field D,
    poisson_package.cpp:112-118): rhs = A exact bit-equal to the oracle's, the solution
    against the oracle's solve with the same faces, and the exact solution recovered."""
    print("case poisson_gmg deck, face field D", flush=True)
    m, om = deck(face_D=True)
    md = m.mesh_data("base")
    Df = deck_face_D(m, 10.0, 1.0)
    md.upload("D", Df)
    eq = PoissonEquation(D_field="D")
    c = centres(m)
    exact = -np.exp(-10.0 * (c[:, 0] ** 2 + c[:, 1] ** 2))
    md.upload("u", to_bricks(m, exact, fill=np.nan))
    pa.exchange_fused(md)
    eq.Ax(md, "u", "rhs")
    rhs = from_bricks(m, md.download("rhs"))
    ref_rhs = om.poisson_apply(exact, Df=Df)
    assert np.array_equal(rhs, ref_rhs), np.max(np.abs(rhs - ref_rhs))
    md.upload("u", np.zeros(m.field_shape("u")))
    p = MeshMGParams(max_iters=60, residual_tolerance=1e-11, precision="fp64",
                     smoother="SRJ2")
    s = MeshMGSolver(m, "u", "rhs", p, eq)
    s.setup(md)
    it, res = s.solve(md)
    print(f"face D: {it} V-cycles, rms residual {res:.3e}, history {s.history()[:6]}",
          flush=True)
    assert res <= 1e-11
    u = from_bricks(m, md.download("u"))
    assert rms(ref_rhs - om.poisson_apply(u, Df=Df)) <= 1.01e-11
    xo, _, _ = om.poisson_solve(ref_rhs, Df=Df, tol=1e-13)
    assert rms(u - xo) <= 1e-10 * rms(xo)
    assert rms(u - exact) <= 1e-10


def test_cpp_deck_driver_with_face_field_D(tmp_path):
    """The regression deck with interior_D = 10: the driver builds the face field D as the
    problem generator does; its rhs and solution match the oracle's with the same faces."""
    print("case cpp deck driver, interior_D = 10", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    text = open(os.path.join(here, "golden", "parthinput.poisson")).read()
    assert "interior_D = 1.0" in text
    deck_file = tmp_path / "parthinput.poisson_D10"
    deck_file.write_text(text.replace("interior_D = 1.0", "interior_D = 10.0")
                         .replace("max_iterations = 15", "max_iterations = 60"))
    out = str(tmp_path / "deckD.bin")
    r = subprocess.run([os.path.join(here, "cpp", "poisson_deck_driver"), str(deck_file), out],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)
    raw = np.fromfile(out, dtype=np.float64)
    res, err, its, nb, per = raw[:5]
    n = int(nb) * int(per)
    u, rhs, exact = (raw[5 + q * n:5 + (q + 1) * n] for q in range(3))
    m, om = deck()
    Df = deck_face_D(m, 10.0, 1.0)
    assert np.array_equal(rhs, om.poisson_apply(exact, Df=Df))
    assert res <= 1e-12 and err <= 1e-11
    xo, _, _ = om.poisson_solve(rhs, Df=Df, tol=1e-13)
    assert rms(u - xo) <= 1e-11


@pytest.mark.parametrize("case", ["deck_face_D", "sphere_3d"])
def test_graph_replay_is_bit_identical_to_eager(case):
    """V-cycles replayed from the captured hipGraph against every V-cycle launched eagerly: the
    same iteration count, residual history and solution bit for bit, over two solves of one
    solver whose right-hand side (and face field D) change between them, so the replays read
    the current level data rather than what the capture saw."""
    print(f"case V-cycle graph replay vs eager, {case}", flush=True)
    if case == "deck_face_D":
        m, _ = deck(face_D=True)
        eq = PoissonEquation(D_field="D")
    else:
        m, _ = make_meshes((32, 32, 32), (8, 8, 8), (-0.5,) * 3, (0.5,) * 3,
                           spheres=[((0.1, -0.05, 0.0), 0.15, 1)])
        eq = PoissonEquation()
    nd = 2 if case == "deck_face_D" else 3
    md = m.mesh_data("base")
    c = centres(m)
    p = MeshMGParams(max_iters=40, residual_tolerance=1e-10, precision="fp64", smoother="SRJ2")
    runs = {}
    for graph in (False, True):
        s = MeshMGSolver(m, "u", "rhs", p, eq)
        s.set_graph(graph)
        out = []
        for seed, interior_D in ((5, 10.0), (6, 4.0)):
            rng = np.random.default_rng(seed)
            b = np.zeros(len(c))
            for _ in range(3):
                x0 = rng.uniform(-0.3, 0.3, 3)
                x0[nd:] = c[0, nd:]  # the blob's centre in the mesh's plane (2-D)
                b += rng.choice([-1.0, 1.0]) * np.exp(-np.sum((c - x0) ** 2, axis=1) / 0.02)
            md.upload("rhs", to_bricks(m, b, name="rhs"))
            md.upload("u", np.zeros(m.field_shape("u")))
            if case == "deck_face_D":
                md.upload("D", deck_face_D(m, interior_D, 1.0))
            s.setup(md)
            it, res = s.solve(md)
            print(f"{case} graph={graph} seed {seed}: {it} V-cycles, rms residual {res:.3e}",
                  flush=True)
            assert res <= 1e-10
            out.append((it, s.history(), md.download("u")))
        runs[graph] = out
    for (ia, ha, ua), (ib, hb, ub) in zip(runs[False], runs[True]):
        assert ia == ib and ia >= 3
        assert np.array_equal(ha, hb)
        assert np.array_equal(ua, ub)
    # the two solves differ: the second did not replay the first's data
    assert not np.array_equal(runs[True][0][2], runs[True][1][2])
