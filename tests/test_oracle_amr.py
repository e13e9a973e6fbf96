"""Pinning tests of the oracle's adaptive remesh (CPU only).

Each check is fixed by something other than the oracle's own formulas: the reference's forest
unit test (block counts after refine / derefine), closed forms of the derivative criteria,
the rules of MeshRefinement::SetRefinement, exact transfer of linear fields (minmod
prolongation and volume-average restriction reproduce them), 2:1 nesting, and conservation of
the Burgers totals across remeshes.
"""
import numpy as np
import pytest

import oracle_lib as O


def derefine_all_possible(m):
    """reference tst/unit/test_forest.cpp:32-49 (DerefineAllPossibleLocations)."""
    cand = []
    for level, l in m.locs():
        if level == 0:
            continue
        p = (level - 1, tuple(x // 2 for x in l))
        kids = [(level, (2 * p[1][0] + i, 2 * p[1][1] + j, 0)) for j in (0, 1) for i in (0, 1)]
        if all(m.is_leaf(lv, ll) for lv, ll in kids):
            cand.append(p)
    cand.sort(key=lambda p: -p[0])
    for lv, ll in cand:
        m.tree_derefine(lv, ll)


def test_forest_refine_derefine_counts():
    # tst/unit/test_forest.cpp:130-165: a periodic 2-D forest of two single-block trees;
    # refining the first block three times gives 5, 11, 23 blocks under proper nesting, and
    # derefining everything possible walks back through 11, 5, 2.
    m = O.OracleMesh((16, 8, 1), (8, 8, 1), 2, periodic=(1, 1, 0))
    assert m.nblocks == 2
    for want in (5, 11, 23):
        level, l = m.locs()[0]
        m.tree_refine(level, l)
        assert m.nblocks == want
    for want in (11, 5, 2):
        derefine_all_possible(m)
        assert m.nblocks == want


@pytest.mark.parametrize("nx,bnx,want", [
    ((128, 128, 128), (16, 16, 16), 3),   # the VIBE benchmark mesh: one tree of 8^3 blocks
    ((64, 64, 64), (16, 16, 16), 2),
    ((64, 32, 32), (16, 16, 16), 1),      # 4x2x2 blocks: 2x1x1 trees of 2^3
    ((48, 48, 48), (16, 16, 16), 0),      # 3^3 single-block trees
    ((64, 64, 1), (16, 16, 1), 0),        # an inactive dimension
])
def test_root_level(nx, bnx, want):
    m = O.OracleMesh(nx, bnx, 2, periodic=(1, 1, 1))
    assert O.OracleAmr(m, 2).root_level == want


def _index_field(m, f):
    """Field whose value at cell (k, j, i) of every block is f(k, j, i) (ghosts included)."""
    shp = m.field_shape(1)
    k, j, i = np.meshgrid(np.arange(shp[2]), np.arange(shp[3]), np.arange(shp[4]),
                          indexing="ij")
    return np.broadcast_to(f(k, j, i).astype(np.float64), shp).copy()


def test_first_derivative_closed_form():
    # q = 100 + i: the x measure 0.5 |q(i+1) - q(i-1)| / |q| = 1 / (100 + i) peaks at the
    # first interior cell i = g; y and z differences vanish.
    m = O.OracleMesh((32, 32, 32), (16, 16, 16), 4)
    a = O.OracleAmr(m, 2)
    q = _index_field(m, lambda k, j, i: 100.0 + i)
    assert np.all(a.derivative_max(q, method=1) == 1.0 / 104.0)
    q = _index_field(m, lambda k, j, i: 100.0 + j + 2.0 * k)   # z difference dominates
    assert np.all(a.derivative_max(q, method=1) == 2.0 / 112.0)   # at j = k = g: q = 112


def test_second_derivative_closed_form():
    # q = 100 + i^2: qavg = q + 1, measure 1 / (|q + 1| + |q|) = 1 / (201 + 2 i^2), peak at i = g
    m = O.OracleMesh((32, 32, 32), (16, 16, 16), 4)
    a = O.OracleAmr(m, 2)
    q = _index_field(m, lambda k, j, i: 100.0 + i * i)
    assert np.all(a.derivative_max(q, method=2) == 1.0 / (201.0 + 2 * 16))


def test_set_refinement_rules():
    # mesh_refinement.cpp:85-117 on a mesh with a refined corner: root blocks never derefine,
    # blocks at the maximum level never refine, a derefine needs `threshold` consecutive
    # derefine tags and no finer neighbour, any other tag resets the count.
    m = O.OracleMesh((64, 64, 64), (16, 16, 16), 4, boxes=[((-0.5, -0.25, -0.5, -0.25, -0.5, -0.25), 1)])
    a = O.OracleAmr(m, numlevel=2, deref_threshold=3)
    root = a.root_level
    locs = m.locs()
    base = next(g for g, (lv, _) in enumerate(locs) if lv == 0)
    fine = [g for g, (lv, _) in enumerate(locs) if lv == 1]
    a.set_refinement(base, -1)
    f, c = a.flags()
    assert (f[base], c[base]) == (0, 0)
    a.set_refinement(base, 1)
    assert a.flags()[0][base] == 1
    # the level-1 blocks are at the maximum level (numlevel 2): refine -> 0
    a.set_refinement(fine[0], 1)
    assert a.flags()[0][fine[0]] == 0
    for n in (1, 2):
        a.set_refinement(fine[0], -1)
        f, c = a.flags()
        assert (f[fine[0]], c[fine[0]]) == (0, n)
    a.set_refinement(fine[0], -1)
    f, c = a.flags()
    assert (f[fine[0]], c[fine[0]]) == (-1, 3)
    a.set_refinement(fine[0], 0)
    f, c = a.flags()
    assert (f[fine[0]], c[fine[0]]) == (0, 0)
    assert root == 2


def test_criterion_level_uses_logical_levels():
    # The criterion's max_level (numlevel) is compared with the logical level root_level +
    # level (refinement_package.cpp:78): on the VIBE mesh (root level 3, numlevel 2) a steep
    # field never asks for refinement; on a mesh with root level 0 it does.
    steep = lambda k, j, i: 1.0 + 5.0 * (i % 4 >= 2)
    for nx, want in (((128, 128, 128), 0), ((48, 48, 48), 1)):
        m = O.OracleMesh(nx, (16, 16, 16), 4)
        a = O.OracleAmr(m, 2, criteria=[dict(method=1, comp=0, refine_tol=0.5, derefine_tol=0.2)])
        a.tag(_index_field(m, steep))
        assert set(a.flags()[0].tolist()) == {want}


def _linear(m, coef=(1.0, 2.0, 3.0, 5.0), ghosts=True):
    """U = c0 + c1 x + c2 y + c3 z at cell centres of every block brick."""
    nb, (e0, e1, e2) = m.nblocks, m.ent
    out = np.zeros((nb, 1, e2, e1, e0))
    for g in range(nb):
        b = m.block(g)
        dx = [(b["xmax"][d] - b["xmin"][d]) / m.n[d] for d in range(3)]
        xc = [b["xmin"][d] + (np.arange(m.ent[d]) - m.ng[d] + 0.5) * dx[d] for d in range(3)]
        z, y, x = np.meshgrid(xc[2], xc[1], xc[0], indexing="ij")
        out[g, 0] = coef[0] + coef[1] * x + coef[2] * y + coef[3] * z
    return out


def _interior(m, a):
    g, n = m.ng, m.n
    return a[:, :, g[2]:g[2] + n[2], g[1]:g[1] + n[1], g[0]:g[0] + n[0]]


def test_transfer_reproduces_linear_fields():
    # Refine two corner blocks and derefine a complete sibling group in one update. Minmod
    # prolongation and volume-average restriction reproduce a linear field exactly (dyadic
    # coordinates), so moved and refined blocks hold the linear field over their whole brick
    # and derefined blocks over their interior (their ghosts start at zero).
    m = O.OracleMesh((64, 64, 64), (8, 8, 8), 4, xmin=(0, 0, 0), xmax=(1, 1, 1),
                     boxes=[((0.0, 0.25, 0.0, 0.25, 0.0, 0.25), 1)])
    a = O.OracleAmr(m, numlevel=3, deref_threshold=1)
    locs = m.locs()
    old = _linear(m)
    # refine two level-0 blocks away from the refined corner, derefine the level-1 group
    refine = [g for g, (lv, l) in enumerate(locs) if lv == 0 and l in ((6, 6, 6), (5, 6, 6))]
    group = [g for g, (lv, l) in enumerate(locs) if lv == 1 and max(l) <= 1]
    assert len(refine) == 2 and len(group) == 8
    for g in range(m.nblocks):
        a.set_refinement(g, 1 if g in refine else (-1 if g in group else 0))
    nb, nnew, ndel = a.update()
    assert (nnew, ndel) == (14, 7) and nb == m.nblocks + 7
    new = a.transfer(old)
    a.commit()
    want = _linear(m)
    newlocs = m.locs()
    for g, (lv, l) in enumerate(newlocs):
        if (lv, l) in locs:                      # moved whole
            np.testing.assert_array_equal(new[g], want[g])
        elif lv == 1:                            # refined: whole brick prolongated
            np.testing.assert_array_equal(new[g], want[g])
        else:                                    # derefined: interior restricted
            assert lv == 0 and l == (0, 0, 0)
            np.testing.assert_array_equal(_interior(m, new)[g], _interior(m, want)[g])
            ghost = new[g].copy()
            gi = m.ng
            ghost[:, gi[2]:-gi[2], gi[1]:-gi[1], gi[0]:-gi[0]] = 0
            assert not ghost.any()


def test_refine_then_derefine_round_trip():
    m = O.OracleMesh((32, 32, 32), (8, 8, 8), 2, xmin=(0, 0, 0), xmax=(1, 1, 1))
    a = O.OracleAmr(m, numlevel=3, deref_threshold=1)
    u0 = _linear(m, coef=(0.5, -1.0, 0.25, 2.0))
    for g in range(m.nblocks):
        a.set_refinement(g, 1 if g in (5, 6) else 0)
    nb, nnew, ndel = a.update()
    u1 = a.transfer(u0)
    a.commit()
    assert m.nblocks == nb and ndel == 0
    lv = np.array([l for l, _ in m.locs()])
    for g in range(m.nblocks):
        a.set_refinement(g, -1 if lv[g] == 1 else 0)
    nb2, nnew2, ndel2 = a.update()
    u2 = a.transfer(u1)
    a.commit()
    assert nb2 == 64 and nnew2 == 0 and ndel2 == nnew
    np.testing.assert_array_equal(_interior(m, u2), _interior(m, u0))


def _levels_nested(m):
    lv = {g: m.block(g)["level"] for g in range(m.nblocks)}
    return all(abs(lv[n["gid"]] - lv[g]) <= 1 for g in range(m.nblocks) for n in m.neighbors(g))


def test_random_updates_keep_nesting_and_constants():
    rng = np.random.default_rng(7)
    m = O.OracleMesh((32, 32, 32), (8, 8, 8), 2)
    a = O.OracleAmr(m, numlevel=4, deref_threshold=1)
    u = np.full(m.field_shape(2), 1.25)
    for step in range(6):
        for g in range(m.nblocks):
            a.set_refinement(g, int(rng.choice([-1, 0, 1], p=[0.45, 0.35, 0.2])))
        nb, nnew, ndel = a.update()
        if nb == 0:
            continue
        u = a.transfer(u)
        a.commit()
        assert _levels_nested(m)
        assert np.all(_interior(m, u) == 1.25)
        m.exchange_bc([u], coarse=[np.zeros(m.coarse_shape(2))])
        assert np.all(u == 1.25)


def test_adaptive_burgers_conserves_totals():
    # Burgers with flux correction on a periodic domain conserves sum(U dV); restriction
    # averages and minmod prolongation preserves cell averages, so remeshing keeps the totals.
    m = O.OracleMesh((32, 32, 32), (8, 8, 8), 4)
    a = O.OracleAmr(m, numlevel=3, deref_threshold=2,
                    criteria=[dict(method=1, comp=3, refine_tol=0.05, derefine_tol=0.01)])
    b = O.OracleBurgers(m, num_scalars=2, cfl=0.4)
    O.lib().orc_burgers_pgen_sinusoid(b.h, 11)
    O.lib().orc_burgers_initialize(b.h)
    t0 = b.totals()
    r = a.burgers_cycles(b, 4)
    assert r >= 1 and m.nblocks != 64
    assert _levels_nested(m)
    np.testing.assert_allclose(b.totals(), t0, rtol=1e-12, atol=1e-12)


def test_criteria_combine_by_maximum():
    # CheckAllRefinement keeps the largest recommendation over the criteria: comp 0 = 100 + i
    # has a first-derivative maximum of 1/(100 + g), comp 1 = 100 + i^2 a second-derivative
    # maximum of 1/(201 + 2 g^2); the tolerances below put each criterion on either side.
    m = O.OracleMesh((48, 48, 48), (16, 16, 16), 4, boxes=[((-0.5, -0.2, -0.5, -0.2, -0.5, -0.2), 1)])
    lv = np.array([l for l, _ in m.locs()])
    q = np.concatenate([_index_field(m, lambda k, j, i: 100.0 + i),
                        _index_field(m, lambda k, j, i: 100.0 + i * i)], axis=1)
    d1, d2 = 1.0 / 104.0, 1.0 / (201.0 + 2 * 16)
    second = dict(method=2, comp=1, refine_tol=2 * d2, derefine_tol=1.5 * d2)   # derefine
    for first, want in ((dict(refine_tol=0.5 * d1, derefine_tol=0.1 * d1), 1),    # refine
                        (dict(refine_tol=2 * d1, derefine_tol=0.5 * d1), 0),      # keep
                        (dict(refine_tol=4 * d1, derefine_tol=2 * d1), -1)):      # derefine
        a = O.OracleAmr(m, numlevel=2, deref_threshold=1,
                        criteria=[dict(method=1, comp=0, **first), second])
        a.tag(q)
        f = a.flags()[0]
        if want == 1:      # base blocks refine; level-1 blocks are at the maximum level
            assert set(f[lv == 0]) == {1} and set(f[lv == 1]) == {0}
        elif want == 0:
            assert not f.any()
        else:              # root blocks never derefine; level-1 blocks without finer neighbours do
            assert set(f[lv == 0]) == {0} and set(f[lv == 1]) == {-1}
