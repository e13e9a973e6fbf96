"""Pins of the oracle's leaf-grid Poisson problem on block meshes (oracle/orc_poisson_mesh.hpp),
the checker of the multigrid solver on refined meshes.

* On a uniform mesh the leaf operator is the 7-point operator of the level-array oracle
  (orc_gmg.hpp, pinned in test_oracle_gmg.py by a manufactured solution), bit for bit.
* On a refined mesh a linear field with dyadic coefficients is reproduced exactly by the
  exchange (restriction averages, linear prolongation), so A u vanishes on every block whose
  stencils stay inside the domain; a wrong slope, a dropped flux or a wrong cell width fails.
* diagonal_alpha enters as -alpha u.
* The BiCGSTAB solve returns a field whose residual, recomputed with the operator, is below
  the tolerance, and recovers a known solution from rhs = A exact (the reference's
  use_exact_rhs mode, example/poisson_gmg/poisson_driver.cpp:87-95).
"""
import numpy as np

from oracle_lib import OracleGmg, OracleMesh


def centres(om):
    n = om.n
    out = []
    for gid in range(om.nblocks):
        b = om.block(gid)
        ax = [b["xmin"][d] + (np.arange(n[d]) + 0.5) * (b["xmax"][d] - b["xmin"][d]) / n[d]
              for d in range(3)]
        z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        out.append(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1))
    return np.concatenate(out)


def test_uniform_leaf_operator_is_the_level_operator():
    N, B = 16, 8
    om = OracleMesh((N,) * 3, (B,) * 3, 2, xmin=(0, 0, 0), xmax=(1, 1, 1), periodic=(0, 0, 0),
                    bcs=("fixed_face",) * 6)
    g = np.random.default_rng(5).uniform(-1, 1, (N, N, N))
    x = np.concatenate([g[l[2] * B:(l[2] + 1) * B, l[1] * B:(l[1] + 1) * B,
                          l[0] * B:(l[0] + 1) * B].ravel()
                        for l in (om.block(b)["l"] for b in range(om.nblocks))])
    got = om.poisson_apply(x)
    ref = OracleGmg(N).apply(g)
    exp = np.concatenate([ref[l[2] * B:(l[2] + 1) * B, l[1] * B:(l[1] + 1) * B,
                            l[0] * B:(l[0] + 1) * B].ravel()
                          for l in (om.block(b)["l"] for b in range(om.nblocks))])
    assert np.array_equal(got, exp)


def test_refined_linear_field_is_harmonic_inside():
    om = OracleMesh((32,) * 3, (8,) * 3, 2, periodic=(1, 1, 1),
                    spheres=[((0.0, 0.0, 0.0), 0.2, 1)])
    assert om.multilevel
    c = centres(om)
    u = 1.0 + 0.5 * c[:, 0] - 0.25 * c[:, 1] + 0.125 * c[:, 2]
    au = om.poisson_apply(u)
    per = 8 ** 3
    checked = 0
    for gid in range(om.nblocks):
        b = om.block(gid)
        if min(b["xmin"]) <= -0.5 + 1e-12 or max(b["xmax"]) >= 0.5 - 1e-12:
            continue  # periodic wrap: the field is not periodic
        assert np.max(np.abs(au[gid * per:(gid + 1) * per])) <= 1e-11
        checked += 1
    levels = {om.block(g)["level"] for g in range(om.nblocks)}
    assert checked >= 16 and levels == {0, 1}


def test_alpha_enters_as_minus_alpha_u():
    om = OracleMesh((64, 64, 1), (32, 32, 1), 2, xmin=(-1, -1, 0), xmax=(1, 1, 1),
                    periodic=(0, 0, 1), bcs=("fixed_face",) * 4,
                    boxes=[((-1, -0.75, -1, -0.75, 0, 1), 3)])
    x = np.random.default_rng(2).uniform(-1, 1, om.nblocks * 32 * 32)
    a0, a1 = om.poisson_apply(x), om.poisson_apply(x, alpha=0.5)
    assert np.allclose(a1, a0 - 0.5 * x, rtol=0, atol=1e-10 * np.max(np.abs(a0)))
    assert np.allclose(om.poisson_apply(x, D=2.0), 2.0 * a0, rtol=1e-13, atol=1e-12)


def test_solve_recovers_exact_solution_on_the_deck():
    om = OracleMesh((64, 64, 1), (32, 32, 1), 2, xmin=(-1, -1, 0), xmax=(1, 1, 1),
                    periodic=(0, 0, 1), bcs=("fixed_face",) * 4,
                    boxes=[((-1, -0.75, -1, -0.75, 0, 1), 3)])
    assert om.nblocks == 13
    c = centres(om)
    exact = -np.exp(-10.0 * (c[:, 0] ** 2 + c[:, 1] ** 2))
    rhs = om.poisson_apply(exact)
    x, it, rms = om.poisson_solve(rhs, tol=1e-13)
    assert it > 0 and rms <= 1e-13
    assert np.sqrt(np.mean((rhs - om.poisson_apply(x)) ** 2)) <= 1e-13
    assert np.sqrt(np.mean((x - exact) ** 2)) <= 1e-12


def test_quadratic_field_on_anisotropic_cells():
    # cell widths 1/32, 1/64, 1/16: A u = 2 (1 + 3 + 0.5) for u = x^2 + 3 y^2 + 0.5 z^2 at every
    # cell whose 7-point stencil stays off the periodic wrap; a cell width used in the wrong
    # direction or a dropped face flux changes the value
    om = OracleMesh((32, 32, 32), (8, 8, 8), 2, xmin=(-0.5, -0.25, -1.0),
                    xmax=(0.5, 0.25, 1.0), periodic=(1, 1, 1))
    c = centres(om)
    u = c[:, 0] ** 2 + 3.0 * c[:, 1] ** 2 + 0.5 * c[:, 2] ** 2
    au = om.poisson_apply(u)
    lo, hi = np.array([-0.5, -0.25, -1.0]), np.array([0.5, 0.25, 1.0])
    h = (hi - lo) / 32
    inside = np.all((c - h > lo) & (c + h < hi), axis=1)
    assert inside.sum() == 30 ** 3
    assert np.max(np.abs(au[inside] - 9.0)) <= 1e-10


def face_field(om, fn):
    """[block][3] bricks of D on each cell's lower faces: component d = fn(d, face centre)."""
    n, g = om.n, om.ng
    e = tuple(n[d] + 2 * g[d] for d in range(3))
    out = np.zeros((om.nblocks, 3) + (e[2], e[1], e[0]))
    for gid in range(om.nblocks):
        b = om.block(gid)
        dx = [(b["xmax"][d] - b["xmin"][d]) / n[d] for d in range(3)]
        cc = [b["xmin"][d] + (np.arange(e[d]) - g[d] + 0.5) * dx[d] for d in range(3)]
        for d in range(3):
            ax = [a.copy() for a in cc]
            ax[d] = ax[d] - 0.5 * dx[d]  # the lower face along d
            z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
            out[gid, d] = fn(d, x, y, z)
    return out


def test_face_field_D_constant_equals_constant_D():
    om = OracleMesh((64, 64, 1), (32, 32, 1), 2, xmin=(-1, -1, 0), xmax=(1, 1, 1),
                    periodic=(0, 0, 1), bcs=("fixed_face",) * 4,
                    boxes=[((-1, -0.75, -1, -0.75, 0, 1), 3)])
    x = np.random.default_rng(4).uniform(-1, 1, om.nblocks * 32 * 32)
    Df = face_field(om, lambda d, a, b, c: np.full(a.shape, 2.5))
    assert np.array_equal(om.poisson_apply(x, Df=Df), om.poisson_apply(x, D=2.5))


def test_face_field_D_flux_continuity():
    # D = 1 on faces with x < 0 and 4 on faces with x > 0 (the face at x = 0 included in the
    # second): a field whose x increments are c / D(face) carries the same flux through every
    # x face, so A u = 0 wherever the stencil stays off the periodic wrap; D read at the cell
    # instead of its faces breaks this at the jump
    om = OracleMesh((32, 8, 8), (8, 8, 8), 2, xmin=(-1, 0, 0), xmax=(1, 0.5, 0.5),
                    periodic=(1, 1, 1))
    Df = face_field(om, lambda d, x, y, z: np.where((d == 0) & (x >= 0), 4.0, 1.0))
    xs = -1.0 + (np.arange(32) + 0.5) / 16
    faces = xs - 1.0 / 32
    Dx = np.where(faces >= 0, 4.0, 1.0)
    ug = np.cumsum(0.25 / Dx)  # u(i) - u(i-1) = 0.25 / D(i - 1/2)
    n = om.n
    u = np.concatenate([np.broadcast_to(ug[om.block(b)["l"][0] * 8:(om.block(b)["l"][0] + 1) * 8],
                                        (n[2], n[1], n[0])).ravel() for b in range(om.nblocks)])
    au = om.poisson_apply(u, Df=Df)
    per = 8 ** 3
    for gid in range(om.nblocks):
        lx = om.block(gid)["l"][0]
        a = au[gid * per:(gid + 1) * per].reshape(8, 8, 8)
        lo, hi = (1 if lx == 0 else 0), (7 if lx == 3 else 8)
        assert np.max(np.abs(a[:, :, lo:hi])) <= 1e-12, gid
    # the same field with D read at the cell centre is not harmonic at the jump
    Dc = face_field(om, lambda d, x, y, z: np.where((d == 0) & (x + 1.0 / 32 >= 0), 4.0, 1.0))
    assert np.max(np.abs(om.poisson_apply(u, Df=Dc))) > 1.0


def test_face_field_D_solve_recovers_exact_solution():
    om = OracleMesh((64, 64, 1), (32, 32, 1), 2, xmin=(-1, -1, 0), xmax=(1, 1, 1),
                    periodic=(0, 0, 1), bcs=("fixed_face",) * 4,
                    boxes=[((-1, -0.75, -1, -0.75, 0, 1), 3)])
    Df = face_field(om, lambda d, x, y, z: np.where((np.abs(x + 0.5) < 0.25) & (np.abs(y) < 0.5),
                                                    10.0, 1.0))
    c = centres(om)
    exact = -np.exp(-10.0 * (c[:, 0] ** 2 + c[:, 1] ** 2))
    rhs = om.poisson_apply(exact, Df=Df)
    x, it, rms = om.poisson_solve(rhs, Df=Df, tol=1e-12)
    assert it > 0 and rms <= 1e-12
    assert np.sqrt(np.mean((rhs - om.poisson_apply(x, Df=Df)) ** 2)) <= 1e-12
    assert np.sqrt(np.mean((x - exact) ** 2)) <= 1e-11
