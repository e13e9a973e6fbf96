"""Pins the oracle's boundary exchange (mesh topology, index ranges, pack/unpack, restriction,
prolongation) against closed forms, before it is trusted as the checker of the HIP path."""
import numpy as np
import pytest

from meshcheck import cell_centers, global_index_field, interior_mask, linear_field
from oracle_lib import OracleMesh


@pytest.mark.parametrize("nx,bnx,g", [
    ((32, 32, 32), (8, 8, 8), 3),    # 4 blocks per dimension
    ((16, 16, 16), (8, 8, 8), 2),    # 2 blocks: +x and -x neighbour is the same block
    ((8, 8, 8), (8, 8, 8), 2),       # 1 block: every neighbour is the block itself
    ((32, 16, 8), (8, 8, 8), 3),     # uneven block counts
])
def test_single_level_ghosts_equal_wrapped_interior(nx, bnx, g):
    m = OracleMesh(nx, bnx, g)
    exp = global_index_field(m, 2, nx)
    f = np.where(interior_mask(m)[None, None], exp, np.nan)
    m.exchange([f])
    assert np.array_equal(f, exp)


def test_buffer_element_counts_match_halo_volume():
    # cfg-1 shape: 512 blocks of 16^3, g=3, 8 one-component variables
    m = OracleMesh((128, 128, 128), (16, 16, 16), 3)
    used, alloc = m.exchange_counts([1] * 8)
    n, g = 16, 3
    halo = (n + 2 * g) ** 3 - n ** 3
    assert used == 512 * 8 * halo
    # GetBufferSize over-allocates: (n+2g+1 | g+1) per dimension
    per_block = 0
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            for oz in (-1, 0, 1):
                if ox == oy == oz == 0:
                    continue
                s = 1
                for o in (ox, oy, oz):
                    s *= (n + 2 * g + 1) if o == 0 else (g + 1)
                per_block += s
    assert alloc == 512 * 8 * per_block


def test_sphere_refined_mesh_block_count():
    m = OracleMesh((64, 64, 64), (16, 16, 16), 4, spheres=[((0.0, 0.0, 0.0), 0.25, 2)])
    levels = [m.block(g)["level"] for g in range(m.nblocks)]
    assert m.multilevel
    assert levels.count(2) == 512 and levels.count(1) == 448 and m.nblocks == 960


def _check_proper_nesting(m):
    for gid in range(m.nblocks):
        lev = m.block(gid)["level"]
        for nb in m.neighbors(gid):
            assert abs(nb["level"] - lev) <= 1


@pytest.mark.parametrize("prolong,case", [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_multilevel_ghosts_reproduce_linear_field(prolong, case):
    # Box refinement up to two levels deep in the middle of a periodic mesh: every ghost cell,
    # whether filled by a same-level copy, a restricted finer neighbour or a prolongated
    # coarser neighbour, must equal the dyadic linear field at its own centre, exactly.
    if case == 0:
        m = OracleMesh((32, 32, 32), (8, 8, 8), 4,
                       boxes=[((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 2)])
    else:
        m = OracleMesh((64, 64, 64), (8, 8, 8), 2,
                       boxes=[((0.0, 0.2, -0.1, 0.1, 0.1, 0.3), 1),
                              ((-0.2, -0.1, -0.2, -0.1, -0.2, -0.1), 2)])
    assert m.multilevel
    _check_proper_nesting(m)
    coeffs = [(3.0, 2.0, 4.0, -8.0), (-1.0, 0.5, -0.25, 1.0)]
    exp = linear_field(m, 2, coeffs)
    f = np.where(interior_mask(m)[None, None], exp, np.nan)
    cf = np.full(m.coarse_shape(2), np.nan)
    # the linear field wraps at the periodic boundary, so only check blocks whose ghost
    # stencils stay inside the domain when they talk to another level
    m.exchange([f], coarse=[cf], prolong=prolong)
    assert not np.isnan(f).any()
    assert np.array_equal(f, exp)


def test_sphere_mesh_ghosts_are_all_filled():
    m = OracleMesh((64, 64, 64), (16, 16, 16), 4, spheres=[((0.0, 0.0, 0.0), 0.25, 2)])
    _check_proper_nesting(m)
    f = np.where(interior_mask(m)[None, None], 1.0, np.nan) * np.ones(m.field_shape(1))
    cf = np.full(m.coarse_shape(1), np.nan)
    m.exchange([f], coarse=[cf], prolong=1)
    # constant data: every ghost cell filled, restriction and prolongation keep the constant
    assert np.array_equal(f, np.ones_like(f))


def restrict_interior(m, ncomp, data):
    from oracle_lib import dptr, lib
    coarse = np.full(m.coarse_shape(ncomp), np.nan)
    lib().orc_restrict_interior(m.h, ncomp, dptr(data), dptr(coarse))
    return coarse


def brute_force_restriction(m, data):
    """2x2x2 averages of every block interior by direct indexing: coarse interior cell
    (ck, cj, ci) is the mean of fine interior cells 2ck..2ck+1, 2cj..2cj+1, 2ci..2ci+1, summed
    in the pairwise order of pr_ops.hpp:157-162 (off-centred pairs first, j pairs inside i
    pairs inside k halves). Cell volumes are equal powers of two here, so the volume weights
    cancel exactly."""
    g, n, cg, cn = m.ng[0], m.n[0], m.cng[0], m.cn[0]
    f = data[:, :, g:g + n, g:g + n, g:g + n]
    t = {(a, b, c): f[:, :, a::2, b::2, c::2] for a in (0, 1) for b in (0, 1) for c in (0, 1)}
    s = (((t[0, 0, 0] + t[0, 1, 0]) + (t[0, 0, 1] + t[0, 1, 1])) +
         ((t[1, 0, 0] + t[1, 1, 0]) + (t[1, 0, 1] + t[1, 1, 1])))
    return s / 8.0, (slice(None), slice(None), slice(cg, cg + cn), slice(cg, cg + cn),
                     slice(cg, cg + cn))


@pytest.mark.parametrize("dyadic", [True, False])
def test_restriction_equals_brute_force_average(dyadic):
    # cfg-2 block shape (16^3, g = 4) on a two-level mesh, 3 components: bit-exact against the
    # direct 2x2x2 average; with dyadic data (multiples of 2^-12 in [-4, 4]) the sum is exact in
    # any order, so the order-free numpy mean must match too
    m = OracleMesh((32, 32, 32), (16, 16, 16), 4, boxes=[((-0.2, 0.2, -0.2, 0.2, -0.2, 0.2), 1)])
    r = np.random.default_rng(5)
    shape = m.field_shape(3)
    data = (np.round(r.uniform(-4, 4, shape) * 4096) / 4096) if dyadic else r.normal(size=shape)
    got = restrict_interior(m, 3, data)
    exp, sl = brute_force_restriction(m, data)
    assert np.array_equal(got[sl], exp)
    if dyadic:
        g, n = m.ng[0], m.n[0]
        f = data[:, :, g:g + n, g:g + n, g:g + n]
        mean = f.reshape(f.shape[0], 3, n // 2, 2, n // 2, 2, n // 2, 2).mean(axis=(3, 5, 7))
        assert np.array_equal(got[sl], mean)
    else:
        # the pairwise order is mirror-symmetric: reversing the fine cells along every axis
        # reverses the coarse cells bit for bit
        rev = np.ascontiguousarray(data[:, :, ::-1, ::-1, ::-1])
        got2 = restrict_interior(m, 3, rev)
        assert np.array_equal(got2[sl], got[sl][:, :, ::-1, ::-1, ::-1])


@pytest.mark.parametrize("prolong", [1, 2], ids=["minmod", "linear"])
def test_prolongation_of_a_parabola_closed_form(prolong):
    # Interior cells hold the cell averages x^2 + h^2/12 of the parabola x^2 (h the cell
    # width); fine blocks sit in the middle of x with coarser blocks on both x sides. Each
    # fine ghost cell across such an x face comes from prolongating its parent coarse cell
    # (centre c, width h, average C = c^2 + h^2/12): its neighbours' averages differ by
    # (2c - h) h below and (2c + h) h above, so the one-sided slopes are 2c - h and 2c + h.
    # ProlongateSharedLinear puts the child at c -/+ h/4 on C -/+ slope * h/4 with the slope of
    # its own side; ProlongateSharedMinMod uses the minmod of the two for both children. The
    # closed forms differ by h^2/2 on the upper child, so a swapped side or limiter fails.
    m = OracleMesh((32, 8, 8), (8, 8, 8), 4, boxes=[((-0.25, 0.25, -0.5, 0.5, -0.5, 0.5), 1)])
    assert m.multilevel
    f = np.full(m.field_shape(1), np.nan)
    g = m.ng[0]
    for gid in range(m.nblocks):
        x, _, _ = cell_centers(m, gid)
        h = (m.block(gid)["xmax"][0] - m.block(gid)["xmin"][0]) / m.n[0]
        f[gid, 0, g:-g, g:-g, g:-g] = (x[g:-g] ** 2 + h * h / 12.0)[None, None, :]
    cf = np.full(m.coarse_shape(1), np.nan)
    m.exchange([f], coarse=[cf], prolong=prolong)
    checked = 0
    for gid in range(m.nblocks):
        b = m.block(gid)
        if b["level"] != 1:
            continue
        hf = (b["xmax"][0] - b["xmin"][0]) / m.n[0]
        h = 2 * hf
        x, _, _ = cell_centers(m, gid)
        for side, cells in (("lo", range(0, g)), ("hi", range(m.n[0] + g, m.ent[0]))):
            edge = b["xmin"][0] if side == "lo" else b["xmax"][0]
            if abs(edge) != 0.25:  # only the faces toward the coarse blocks
                continue
            for i in cells:
                xf = x[i]
                c = (np.floor(xf / h) + 0.5) * h
                C = c * c + h * h / 12.0
                lower = xf < c
                slope_m, slope_p = 2 * c - h, 2 * c + h
                if prolong == 2:
                    exp = C - slope_m * h / 4 if lower else C + slope_p * h / 4
                else:
                    gc = (np.sign(slope_m) + np.sign(slope_p)) / 2 * min(abs(slope_m),
                                                                         abs(slope_p))
                    exp = C - gc * h / 4 if lower else C + gc * h / 4
                got = f[gid, 0, g:-g, g:-g, i]
                assert np.max(np.abs(got - exp)) < 1e-15, (gid, side, i, got.ravel()[0], exp)
                checked += 1
    assert checked >= 8
