"""Pins the oracle's physical boundary conditions (reference GenericBC Outflow / Reflect,
boundary_conditions_generic.hpp:174-255, applied face by face in BoundaryFace order by
ApplyBoundaryConditionsOnCoarseOrFine, boundary_conditions.cpp:36-58) against closed forms:
after the exchange and the conditions, every cell holds the value of its source cell, found
dimension by dimension (wrap, clamp or mirror), with the vector sign rule."""
import numpy as np
import pytest

from meshcheck import bc_index_field, interior_mask
from oracle_lib import OracleMesh

CASES = [
    # one block: every ghost comes from the conditions (corners compose three faces)
    ((8, 8, 8), (8, 8, 8), 3, ("outflow",) * 6),
    ((8, 8, 8), (8, 8, 8), 3, ("reflect",) * 6),
    ((8, 8, 8), (8, 8, 8), 4, ("reflect", "outflow", "outflow", "reflect", "reflect",
                               "outflow")),
    # several blocks, one periodic dimension: edges and corners mix exchanged and
    # condition-filled ghosts
    ((32, 16, 16), (8, 8, 8), 3, ("reflect", "outflow", "periodic", "periodic", "outflow",
                                  "reflect")),
    ((16, 16, 32), (8, 8, 8), 4, ("outflow", "outflow", "reflect", "reflect", "periodic",
                                  "periodic")),
    # fixed_face (zero face value, the Dirichlet faces of the Poisson solvers): the ghost is
    # the negated mirror cell, and corners crossing two fixed faces get the sign back
    ((8, 8, 8), (8, 8, 8), 3, ("fixed_face",) * 6),
    ((16, 16, 16), (8, 8, 8), 4, ("fixed_face", "outflow", "reflect", "fixed_face",
                                  "periodic", "periodic")),
]


def periodic_flags(bcs):
    return tuple(int(bcs[2 * d] == "periodic") for d in range(3))


@pytest.mark.parametrize("nx,bnx,g,bcs", CASES)
@pytest.mark.parametrize("vector", [False, True])
def test_ghosts_equal_source_cells(nx, bnx, g, bcs, vector):
    m = OracleMesh(nx, bnx, g, periodic=periodic_flags(bcs), bcs=bcs)
    exp = bc_index_field(m, 4, nx, bcs, vector=vector)
    f = np.where(interior_mask(m)[None, None], exp, np.nan)
    m.exchange_bc([f], vector=[vector])
    assert not np.isnan(f).any()
    assert np.array_equal(f, exp)


def test_conditions_alone_touch_only_boundary_ghosts():
    # blocks away from the domain boundary and the periodic dimension keep their ghosts
    bcs = ("outflow", "reflect", "periodic", "periodic", "reflect", "outflow")
    nx = (32, 16, 16)
    m = OracleMesh(nx, (8, 8, 8), 2, periodic=periodic_flags(bcs), bcs=bcs)
    rng = np.random.default_rng(0)
    f = rng.standard_normal(m.field_shape(3))
    before = f.copy()
    m.apply_bc([f], vector=[True])
    g, e = m.ng, m.ent
    for gid in range(m.nblocks):
        b = m.block(gid)
        inner = b["l"][0] not in (0, nx[0] // 8 - 1) and b["l"][2] not in (0, nx[2] // 8 - 1)
        if inner:
            assert np.array_equal(f[gid], before[gid])
    # interiors are never written
    mask = interior_mask(m)
    assert np.array_equal(f[:, :, mask], before[:, :, mask])


def test_unknown_or_periodic_face_rejected():
    with pytest.raises(ValueError):
        OracleMesh((8, 8, 8), (8, 8, 8), 2, periodic=(1, 0, 0),
                   bcs=("outflow", "outflow", "outflow", "outflow", "outflow", "outflow"))


@pytest.mark.parametrize("bcs", [
    ("outflow",) * 6,
    ("reflect", "outflow", "periodic", "periodic", "outflow", "reflect"),
])
def test_multilevel_constant_field_stays_constant(bcs):
    # constant data on a refined mesh with physical faces: the coarse-buffer conditions feed
    # the prolongation next to the boundary, and every ghost must keep the constant
    m = OracleMesh((32, 32, 32), (8, 8, 8), 4, periodic=periodic_flags(bcs), bcs=bcs,
                   boxes=[((-0.5, -0.3, -0.1, 0.1, -0.5, -0.2), 1)])
    assert m.multilevel
    vals = np.array([1.5, -2.0, 3.25])
    full = np.ones(m.field_shape(3)) * vals[None, :, None, None, None]
    f = np.where(interior_mask(m)[None, None], full, np.nan)
    # coarse cells the exchange never writes hold NaN: none may reach the fine data
    cf = np.full(m.coarse_shape(3), np.nan)
    m.exchange_bc([f], coarse=[cf])
    assert np.array_equal(f, full)
