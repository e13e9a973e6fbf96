"""Closed-form expectations for ghost zones, shared by oracle and GPU tests.

These do not restate any algorithm of the reference: they state what the data must be once
the exchange is correct (ghost cell == the value its physical position has in the owning
block), so they pin the oracle and the HIP path independently of each other.
"""
import numpy as np


def cell_centers(mesh_like, gid, coarse=False):
    """Cell-centre coordinates (x, y, z) arrays of block gid over its entire index range,
    using the UniformCartesian formula xmin - g*dx + (i + 0.5)*dx (or its coarse version)."""
    b = mesh_like.block(gid)
    out = []
    for d in range(3):
        n = mesh_like.n[d]
        g = mesh_like.ng[d]
        dx = (b["xmax"][d] - b["xmin"][d]) / n
        x0 = b["xmin"][d] - g * dx
        ent = mesh_like.ent[d]
        if coarse:
            x0 = x0 + g * dx * (1 - 2)
            dx = dx * 2 if g > 0 else dx
            ent = mesh_like.cent[d]
        out.append(x0 + (np.arange(ent) + 0.5) * dx)
    return out


def global_index_field(mesh_like, ncomp, nx):
    """Single-level periodic mesh: value = c*1e7 + global cell id of the (wrapped) cell."""
    shape = (mesh_like.nblocks, ncomp) + tuple(reversed(mesh_like.ent))
    expected = np.empty(shape)
    for gid in range(mesh_like.nblocks):
        b = mesh_like.block(gid)
        assert b["level"] == 0
        idx = []
        for d in range(3):
            g = mesh_like.ng[d]
            gi = b["l"][d] * mesh_like.n[d] + np.arange(mesh_like.ent[d]) - g
            idx.append(np.mod(gi, nx[d]))
        gk, gj, gi = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
        cell = (gk * nx[1] + gj) * nx[0] + gi
        for c in range(ncomp):
            expected[gid, c] = c * 1.0e7 + cell
    return expected


def interior_mask(mesh_like):
    e, g = mesh_like.ent, mesh_like.ng
    m = np.zeros(tuple(reversed(e)), dtype=bool)
    m[g[2]:e[2] - g[2], g[1]:e[1] - g[1], g[0]:e[0] - g[0]] = True
    return m


def linear_field(mesh_like, ncomp, coeffs, domain=((-0.5, 0.5),) * 3):
    """Per-component dyadic linear field f_c = a_c + sum_d b_cd x_d evaluated at the wrapped
    cell centre of every cell (ghosts included)."""
    shape = (mesh_like.nblocks, ncomp) + tuple(reversed(mesh_like.ent))
    out = np.empty(shape)
    for gid in range(mesh_like.nblocks):
        xs = cell_centers(mesh_like, gid)
        wr = []
        for d in range(3):
            lo, hi = domain[d]
            x = xs[d].copy()
            x[x < lo] += hi - lo
            x[x > hi] -= hi - lo
            wr.append(x)
        z, y, x = np.meshgrid(wr[2], wr[1], wr[0], indexing="ij")
        for c in range(ncomp):
            a, b1, b2, b3 = coeffs[c]
            out[gid, c] = a + b1 * x + b2 * y + b3 * z
    return out


def bc_index_field(mesh_like, ncomp, nx, bcs, vector=False):
    """Single-level mesh with physical boundaries: value = sign * (c*1e7 + global cell id) of
    the source cell of every cell. Along each dimension a global index outside the mesh maps
    back inside by its face's condition: periodic wraps, outflow copies the first / last cell,
    reflect and fixed_face mirror across the face. Component c < 3 of a vector field changes
    sign once per reflecting face it crosses along dimension c; every component changes sign
    once per fixed_face face it crosses (ghost = 2 * 0 - mirror cell, a zero face value)."""
    shape = (mesh_like.nblocks, ncomp) + tuple(reversed(mesh_like.ent))
    expected = np.empty(shape)
    for gid in range(mesh_like.nblocks):
        b = mesh_like.block(gid)
        assert b["level"] == 0
        idx, refl, fixed = [], [], []
        for d in range(3):
            g = mesh_like.ng[d]
            gi = b["l"][d] * mesh_like.n[d] + np.arange(mesh_like.ent[d]) - g
            src = gi.copy()
            r = np.zeros(gi.shape, dtype=bool)
            f = np.zeros(gi.shape, dtype=bool)
            lo, hi = gi < 0, gi >= nx[d]
            for mask, face, edge, mirror in ((lo, bcs[2 * d], 0, -1 - gi),
                                             (hi, bcs[2 * d + 1], nx[d] - 1,
                                              2 * nx[d] - 1 - gi)):
                if face == "periodic":
                    src[mask] = np.mod(gi[mask], nx[d])
                elif face == "outflow":
                    src[mask] = edge
                elif face == "fixed_face":
                    src[mask] = mirror[mask]
                    f |= mask
                else:
                    src[mask] = mirror[mask]
                    r |= mask
            idx.append(src)
            refl.append(r)
            fixed.append(f)
        gk, gj, gi = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
        rk, rj, ri = np.meshgrid(refl[2], refl[1], refl[0], indexing="ij")
        fk, fj, fi = np.meshgrid(fixed[2], fixed[1], fixed[0], indexing="ij")
        nfix = fi.astype(int) + fj.astype(int) + fk.astype(int)
        cell = (gk * nx[1] + gj) * nx[0] + gi
        for c in range(ncomp):
            v = c * 1.0e7 + cell.astype(np.float64)
            if vector and c < 3:
                v = np.where((ri, rj, rk)[c], -v, v)
            expected[gid, c] = np.where(nfix % 2 == 1, -v, v)
    return expected
