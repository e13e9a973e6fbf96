"""Pins the oracle's geometric-multigrid Poisson solver against what the mathematics fixes,
before it checks the GPU solver: second-order consistency of the 7-point operator with the
FixedFace-0 boundary on a manufactured solution, symmetry and definiteness, exactness of the
transfer operators on constant and linear data, the level hierarchy of the reference's GMG
grids, and the known-answer solve of the reference's poisson_gmg test (rhs = A u*, solved
back to u*)."""
import numpy as np
import pytest

from oracle_lib import OracleGmg

rng = np.random.default_rng(11)


def centres(N):
    x = (np.arange(N) + 0.5) / N
    return np.meshgrid(x, x, x, indexing="ij")  # (z, y, x) order matches [k][j][i]


def test_hierarchy_halves_while_even():
    assert OracleGmg(256).levels == 9      # 256 ... 1
    assert OracleGmg(192).levels == 7      # 192 96 48 24 12 6 3
    assert OracleGmg(32).levels == 6


def test_operator_is_second_order_on_a_manufactured_solution():
    errs = []
    for N in (16, 32):
        z, y, x = centres(N)
        u = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
        Au = OracleGmg(N).apply(u)
        errs.append(np.max(np.abs(Au + 3 * np.pi ** 2 * u)))
    assert errs[1] < errs[0] / 3.5, errs   # O(h^2): ratio 4 on halving h


def test_operator_symmetric_negative_definite_and_linear():
    g = OracleGmg(16)
    x, y = rng.standard_normal((2, 16, 16, 16))
    Ax, Ay = g.apply(x), g.apply(y)
    assert abs(np.sum(x * Ay) - np.sum(y * Ax)) < 1e-10 * np.sum(np.abs(x * Ay))
    assert np.sum(x * Ax) < 0
    assert np.allclose(g.apply(2.0 * x - 3.0 * y), 2.0 * Ax - 3.0 * Ay, rtol=0, atol=1e-9)


def test_restriction_averages_and_prolongation_is_linear_exact():
    g = OracleGmg(16)
    c = np.full((16, 16, 16), 1.75)
    assert np.array_equal(g.restrict(0, c), np.full((8, 8, 8), 1.75))
    z, y, x = centres(16)
    lin = 0.5 + 2.0 * x - 1.0 * y + 0.25 * z
    zc, yc, xc = centres(8)
    linc = 0.5 + 2.0 * xc - 1.0 * yc + 0.25 * zc
    assert np.allclose(g.restrict(0, lin), linc, atol=1e-14)
    # prolongation of a linear coarse field: exact away from the boundary (whose ghosts follow
    # the Dirichlet condition, not the linear field)
    f = g.prolong_add(0, linc, np.zeros((16, 16, 16)))
    assert np.allclose(f[2:-2, 2:-2, 2:-2], lin[2:-2, 2:-2, 2:-2], atol=1e-13)
    # prolongation of a constant at the boundary sees ghost = -interior
    f = g.prolong_add(0, np.ones((8, 8, 8)), np.zeros((16, 16, 16)))
    assert np.allclose(f[4:-4, 4:-4, 4:-4], 1.0)
    assert np.all(f[0] < 1.0)


@pytest.mark.parametrize("N", [32, 48])
def test_known_answer_solve_recovers_exact_solution(N):
    # the reference poisson_gmg known-answer mode: rhs = A u* for a known u*
    g = OracleGmg(N)
    z, y, x = centres(N)
    ustar = np.sin(np.pi * x) * np.sin(2 * np.pi * y) * z * (1 - z) + 0.1 * x * (1 - x) * y
    b = g.apply(ustar)
    g.set_rhs(b)
    it, hist = g.solve(tol=1e-10, maxit=60)
    assert hist[-1] < 1e-10 and it < 40
    # V-cycle convergence factor of 2+2 weighted Jacobi (omega 2/3): the 3-D smoothing factor
    # of omega-Jacobi is max(|1 - omega/3|, |1 - 2 omega|) = 7/9, so four sweeps bound the
    # two-grid factor near (7/9)^4 = 0.37
    factors = hist[1:] / hist[:-1]
    assert np.max(factors[1:]) < 0.4, factors
    err = np.sqrt(np.mean((g.u() - ustar) ** 2)) / np.sqrt(np.mean(ustar ** 2))
    assert err < 1e-9, err


def test_blob_rhs_is_seeded_and_solves():
    g = OracleGmg(32)
    b1 = g.rhs_blobs(3)
    b2 = OracleGmg(32).rhs_blobs(3)
    assert np.array_equal(b1, b2) and np.max(np.abs(b1)) > 0.1
    it, hist = g.solve(tol=1e-10, maxit=60)
    assert hist[-1] < 1e-10
    # the solution satisfies the discrete equation
    r = g.apply(g.u()) - b1
    assert np.linalg.norm(r) / np.linalg.norm(b1) < 1e-10
