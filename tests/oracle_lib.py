"""ctypes wrapper over oracle/liboracle.so (the CPU restatement used as the parity checker).

Test infrastructure only: imported by tests/, the smoke check and bench.py's CPU-baseline
leg, never by the product package.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

D = C.c_double
I = C.c_int
I64 = C.c_int64
P = C.c_void_p
DP = C.POINTER(C.c_double)


def lib():
    global _LIB
    if _LIB is None:
        # PTH_ORACLE_LIB points the tests at another build of the restatement (a planted
        # mistake, to show that the pins catch it)
        path = os.environ.get("PTH_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        sig = {
            "orc_mesh_new": (P, [C.POINTER(I), C.POINTER(I), I, DP, DP, C.POINTER(I)]),
            "orc_mesh_refine_box": (None, [P, DP, I]),
            "orc_mesh_refine_sphere": (None, [P, DP, D, I]),
            "orc_mesh_finalize": (None, [P]),
            "orc_mesh_free": (None, [P]),
            "orc_mesh_nblocks": (I, [P]),
            "orc_mesh_multilevel": (I, [P]),
            "orc_mesh_block": (None, [P, I, C.POINTER(I64), DP]),
            "orc_mesh_nneighbors": (I, [P, I]),
            "orc_mesh_neighbor": (None, [P, I, I, C.POINTER(I64)]),
            "orc_mesh_shapes": (None, [P, C.POINTER(I)]),
            "orc_mesh_region_boxes": (I, [P, I, C.POINTER(I), I]),
            "orc_exchange_counts": (I64, [P, I, C.POINTER(I), C.POINTER(I64)]),
            "orc_exchange": (None, [P, I, C.POINTER(I), C.POINTER(DP), C.POINTER(DP), I, I]),
            "orc_mesh_set_boundary": (I, [P, I, I]),
            "orc_exchange_bc": (None, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(DP),
                                       C.POINTER(DP), I]),
            "orc_apply_bc": (None, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(DP), I]),
            "orc_restrict_interior": (None, [P, I, DP, DP]),
            "orc_prolongate_cells": (None, [P, I, DP, DP, I, C.POINTER(I), I]),
            "orc_poisson_leaf_apply": (None, [P, D, D, DP, DP, DP]),
            "orc_poisson_leaf_solve": (I, [P, D, D, DP, DP, DP, D, I, DP]),
            "orc_weno5z": (None, [I, DP, DP, DP]),
            "orc_linear": (None, [I, DP, DP, DP]),
            "orc_lr_to_flux": (None, [I, DP, DP]),
            "orc_scalar_flux": (None, [I, DP, DP]),
            "orc_burgers_new": (P, [P, I, D, I]),
            "orc_burgers_free": (None, [P]),
            "orc_burgers_pgen_sinusoid": (None, [P, C.c_uint64]),
            "orc_burgers_pgen_vibe": (None, [P]),
            "orc_burgers_initialize": (None, [P]),
            "orc_burgers_cycles": (I, [P, I]),
            "orc_burgers_set_time_limits": (None, [P, DP, I, I]),
            "orc_burgers_ncycle": (I, [P]),
            "orc_burgers_fluxes": (None, [P]),
            "orc_burgers_fluxdiv": (None, [P]),
            "orc_burgers_derived": (None, [P]),
            "orc_burgers_estimate_dt": (D, [P]),
            "orc_burgers_dt": (D, [P]),
            "orc_burgers_time": (D, [P]),
            "orc_burgers_size": (I64, [P, I]),
            "orc_burgers_get": (None, [P, I, DP]),
            "orc_burgers_set": (None, [P, I, DP]),
            "orc_burgers_mass_history": (D, [P, DP, DP, D]),
            "orc_burgers_set_flux_correction": (None, [P, I]),
            "orc_burgers_totals": (None, [P, DP]),
            "orc_gmg_new": (P, [I]),
            "orc_gmg_free": (None, [P]),
            "orc_gmg_levels": (I, [P]),
            "orc_gmg_set_params": (None, [P, D, I, I]),
            "orc_gmg_rhs_blobs": (None, [P, C.c_uint64]),
            "orc_gmg_set_rhs": (None, [P, DP]),
            "orc_gmg_get_rhs": (None, [P, DP]),
            "orc_gmg_set_u": (None, [P, DP]),
            "orc_gmg_get_u": (None, [P, DP]),
            "orc_gmg_apply": (None, [P, DP, DP]),
            "orc_gmg_solve": (I, [P, D, I]),
            "orc_gmg_history": (I, [P, DP, I]),
            "orc_gmg_restrict": (None, [P, I, DP, DP]),
            "orc_gmg_prolong_add": (None, [P, I, DP, DP]),
            "orc_amr_new": (P, [P, I, I]),
            "orc_amr_free": (None, [P]),
            "orc_amr_add_criterion": (None, [P, I, I, D, D, I]),
            "orc_amr_root_level": (I, [P]),
            "orc_amr_derivative_max": (None, [P, I, I, I, DP, DP]),
            "orc_amr_tag": (None, [P, I, DP]),
            "orc_amr_set_refinement": (None, [P, I, I]),
            "orc_amr_flags": (None, [P, C.POINTER(I), C.POINTER(I)]),
            "orc_amr_update": (I, [P, C.POINTER(I), C.POINTER(I)]),
            "orc_amr_transfer": (None, [P, I, DP, DP]),
            "orc_amr_commit": (None, [P]),
            "orc_amr_burgers_cycles": (I, [P, P, I]),
            "orc_mesh_tree_refine": (I, [P, I, C.POINTER(I64)]),
            "orc_mesh_tree_derefine": (I, [P, I, C.POINTER(I64)]),
            "orc_mesh_is_leaf": (I, [P, I, C.POINTER(I64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def iarr(vals, t=C.c_int):
    return (t * len(vals))(*vals)


BC_TYPES = {"periodic": 0, "outflow": 1, "reflect": 2, "fixed_face": 3}


class OracleMesh:
    """Mesh restatement: uniform base grid, optional static refinement boxes/spheres."""

    def __init__(self, nx, bnx, nghost, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                 periodic=(1, 1, 1), boxes=(), spheres=(), bcs=None):
        """bcs: optional 6 face conditions (inner x1, outer x1, ..., outer x3), each
        "periodic", "outflow" or "reflect"; faces of non-periodic dimensions default to
        outflow."""
        L = lib()
        self.h = L.orc_mesh_new(iarr(nx), iarr(bnx), nghost, (D * 3)(*xmin), (D * 3)(*xmax),
                                iarr(periodic))
        for face, bc in enumerate(bcs or ()):
            if bc == "periodic":
                continue
            if L.orc_mesh_set_boundary(self.h, face, BC_TYPES[bc]) != 0:
                raise ValueError(f"face {face}: boundary {bc!r} not allowed")
        for box, level in boxes:
            L.orc_mesh_refine_box(self.h, (D * 6)(*box), level)
        for center, r, level in spheres:
            L.orc_mesh_refine_sphere(self.h, (D * 3)(*center), r, level)
        L.orc_mesh_finalize(self.h)
        self.refresh()
        s = (I * 12)()
        L.orc_mesh_shapes(self.h, s)
        self.n = tuple(s[0:3])
        self.ng = tuple(s[3:6])
        self.cn = tuple(s[6:9])
        self.cng = tuple(s[9:12])
        self.ent = tuple(self.n[d] + 2 * self.ng[d] for d in range(3))
        self.cent = tuple(self.cn[d] + 2 * self.cng[d] for d in range(3))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_mesh_free(self.h)
            self.h = None

    def poisson_apply(self, x, D=1.0, alpha=0.0, Df=None):
        """Leaf-grid Poisson operator (exchange with linear prolongation and the face
        conditions, then Ax) on interior values in block, then [k][j][i] order; Df: the face
        field D as [block][3] bricks, or None for the constant D."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        out = np.empty_like(x)
        df = None if Df is None else np.ascontiguousarray(Df, dtype=np.float64)
        lib().orc_poisson_leaf_apply(self.h, D, alpha, None if df is None else dptr(df),
                                     dptr(x), dptr(out))
        return out

    def poisson_solve(self, b, x0=None, D=1.0, alpha=0.0, Df=None, tol=1e-14, maxit=5000):
        """BiCGSTAB on the leaf grid to RMS residual tol; returns (x, iterations, rms)."""
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64).ravel()
        df = None if Df is None else np.ascontiguousarray(Df, dtype=np.float64)
        rms = C.c_double(0.0)
        it = lib().orc_poisson_leaf_solve(self.h, D, alpha, None if df is None else dptr(df),
                                          dptr(b), dptr(x), tol, maxit, C.byref(rms))
        return x, it, rms.value

    def refresh(self):
        """Re-read the block count after the tree changed (remesh, direct refine/derefine)."""
        self.nblocks = lib().orc_mesh_nblocks(self.h)
        self.multilevel = bool(lib().orc_mesh_multilevel(self.h))

    def locs(self):
        """(level, (l1, l2, l3)) of every block in gid order."""
        out = []
        for g in range(self.nblocks):
            b = self.block(g)
            out.append((b["level"], b["l"]))
        return out

    def tree_refine(self, level, l):
        n = lib().orc_mesh_tree_refine(self.h, level, iarr(l, I64))
        self.refresh()
        return n

    def tree_derefine(self, level, l):
        n = lib().orc_mesh_tree_derefine(self.h, level, iarr(l, I64))
        self.refresh()
        return n

    def is_leaf(self, level, l):
        return bool(lib().orc_mesh_is_leaf(self.h, level, iarr(l, I64)))

    def block(self, gid):
        out = (I64 * 4)()
        x = (D * 6)()
        lib().orc_mesh_block(self.h, gid, out, x)
        return dict(level=out[0], l=tuple(out[1:4]), xmin=tuple(x[0:3]), xmax=tuple(x[3:6]))

    def neighbors(self, gid):
        res = []
        for n in range(lib().orc_mesh_nneighbors(self.h, gid)):
            out = (I64 * 8)()
            lib().orc_mesh_neighbor(self.h, gid, n, out)
            res.append(dict(gid=out[0], level=out[1], origin=tuple(out[2:5]),
                            off=tuple(out[5:8])))
        return res

    def region_boxes(self, gid):
        buf = (I * (7 * 512))()
        n = lib().orc_mesh_region_boxes(self.h, gid, buf, 512)
        return np.array(buf[:7 * n], dtype=np.int64).reshape(n, 7)

    def field_shape(self, ncomp):
        e = self.ent
        return (self.nblocks, ncomp, e[2], e[1], e[0])

    def coarse_shape(self, ncomp):
        e = self.cent
        return (self.nblocks, ncomp, e[2], e[1], e[0])

    def exchange_counts(self, ncomps):
        alloc = I64()
        used = lib().orc_exchange_counts(self.h, len(ncomps), iarr(ncomps), C.byref(alloc))
        return used, alloc.value

    def exchange(self, fields, coarse=None, prolong=0, nrep=1):
        """In-place Send+Set(+prolongation) over a list of float64 arrays."""
        ncomps = [f.shape[1] for f in fields]
        dps = (DP * len(fields))(*[dptr(f) for f in fields])
        if coarse is None:
            cps = (DP * len(fields))(*[None] * len(fields))
        else:
            cps = (DP * len(fields))(*[dptr(c) for c in coarse])
        lib().orc_exchange(self.h, len(fields), iarr(ncomps), dps, cps, prolong, nrep)

    def exchange_bc(self, fields, coarse=None, vector=None, nrep=1):
        """In-place exchange in library order with physical boundary conditions: send, set,
        coarse-buffer conditions, minmod prolongation (multilevel), conditions."""
        ncomps = [f.shape[1] for f in fields]
        vec = iarr([int(v) for v in (vector or [0] * len(fields))])
        dps = (DP * len(fields))(*[dptr(f) for f in fields])
        if coarse is None:
            cps = (DP * len(fields))(*[None] * len(fields))
        else:
            cps = (DP * len(fields))(*[dptr(c) for c in coarse])
        lib().orc_exchange_bc(self.h, len(fields), iarr(ncomps), vec, dps, cps, nrep)

    def apply_bc(self, arrays, vector=None, coarse=False):
        """In-place physical boundary conditions on data (or coarse) arrays."""
        ncomps = [f.shape[1] for f in arrays]
        vec = iarr([int(v) for v in (vector or [0] * len(arrays))])
        dps = (DP * len(arrays))(*[dptr(f) for f in arrays])
        lib().orc_apply_bc(self.h, len(arrays), iarr(ncomps), vec, dps, int(coarse))


class OracleBurgers:
    def __init__(self, mesh, num_scalars=4, cfl=0.4, recon=0):
        self.mesh = mesh
        self.V = 3 + num_scalars
        self.h = lib().orc_burgers_new(mesh.h, num_scalars, cfl, recon)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_burgers_free(self.h)
            self.h = None

    def get(self, which):
        n = lib().orc_burgers_size(self.h, which)
        out = np.empty(n, dtype=np.float64)
        lib().orc_burgers_get(self.h, which, dptr(out))
        return out

    def set(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == lib().orc_burgers_size(self.h, which)
        lib().orc_burgers_set(self.h, which, dptr(a))

    @property
    def dt(self):
        return lib().orc_burgers_dt(self.h)

    @property
    def time(self):
        return lib().orc_burgers_time(self.h)

    @property
    def ncycle(self):
        return lib().orc_burgers_ncycle(self.h)

    def set_time_limits(self, tlim=math.inf, dt_init=sys.float_info.max, dt_init_force=False,
                        dt_user=sys.float_info.max, dt_force=-sys.float_info.max,
                        dt_floor=sys.float_info.min, dt_ceil=sys.float_info.max, nlim=-1):
        """<parthenon/time> limits (EvolutionDriver defaults); before initialize()."""
        v = np.array([tlim, dt_init, dt_user, dt_force, dt_floor, dt_ceil], dtype=np.float64)
        lib().orc_burgers_set_time_limits(self.h, dptr(v), int(dt_init_force), nlim)

    def cycles(self, n):
        """EvolutionDriver loop: at most n cycles while KeepGoing(); returns cycles run."""
        return lib().orc_burgers_cycles(self.h, n)

    def totals(self):
        out = np.empty(self.V)
        lib().orc_burgers_totals(self.h, dptr(out))
        return out


class OracleAmr:
    """Adaptive-refinement state of an OracleMesh: criteria, per-block flags, remesh."""

    def __init__(self, mesh, numlevel, deref_threshold=10, criteria=()):
        self.mesh = mesh
        self.h = lib().orc_amr_new(mesh.h, numlevel, deref_threshold)
        for c in criteria:
            self.add_criterion(**c)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_amr_free(self.h)
            self.h = None

    def add_criterion(self, method=1, comp=0, refine_tol=0.5, derefine_tol=0.05, max_level=0):
        lib().orc_amr_add_criterion(self.h, method, comp, refine_tol, derefine_tol, max_level)

    @property
    def root_level(self):
        return lib().orc_amr_root_level(self.h)

    def derivative_max(self, data, comp=0, method=1):
        data = np.ascontiguousarray(data, dtype=np.float64)
        out = np.empty(self.mesh.nblocks)
        lib().orc_amr_derivative_max(self.h, method, data.shape[1], comp, dptr(data), dptr(out))
        return out

    def tag(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        lib().orc_amr_tag(self.h, data.shape[1], dptr(data))

    def set_refinement(self, gid, tag):
        lib().orc_amr_set_refinement(self.h, gid, tag)

    def flags(self):
        n = self.mesh.nblocks
        f, c = (I * n)(), (I * n)()
        lib().orc_amr_flags(self.h, f, c)
        return np.array(f[:], dtype=np.int32), np.array(c[:], dtype=np.int32)

    def update(self):
        """UpdateMeshBlockTree: (new block count or 0 if unchanged, nnew, ndel)."""
        nnew, ndel = I(), I()
        nb = lib().orc_amr_update(self.h, C.byref(nnew), C.byref(ndel))
        self.pending = nb
        return nb, nnew.value, ndel.value

    def transfer(self, data):
        """Moves one field of the current mesh onto the pending block list."""
        data = np.ascontiguousarray(data, dtype=np.float64)
        e = self.mesh.ent
        out = np.empty((self.pending, data.shape[1], e[2], e[1], e[0]))
        lib().orc_amr_transfer(self.h, data.shape[1], dptr(data), dptr(out))
        return out

    def commit(self):
        lib().orc_amr_commit(self.h)
        self.mesh.refresh()

    def burgers_cycles(self, burgers, n):
        """n adaptive Burgers cycles (tag + remesh after each); returns remesh count."""
        r = lib().orc_amr_burgers_cycles(self.h, burgers.h, n)
        self.mesh.refresh()
        return r


class OracleGmg:
    """FP64 geometric-multigrid Poisson solver on the uniform N^3 cell-centred grid of [0,1]^3
    (FixedFace 0 boundaries); arrays are interior (N, N, N) in [k][j][i] order."""

    def __init__(self, N, omega=2.0 / 3.0, pre=2, post=2):
        self.N = N
        self.h = lib().orc_gmg_new(N)
        lib().orc_gmg_set_params(self.h, omega, pre, post)
        self.levels = lib().orc_gmg_levels(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_gmg_free(self.h)
            self.h = None

    def _arr(self, n=None):
        n = self.N if n is None else n
        return np.zeros((n, n, n))

    def rhs_blobs(self, seed):
        lib().orc_gmg_rhs_blobs(self.h, seed)
        out = self._arr()
        lib().orc_gmg_get_rhs(self.h, dptr(out))
        return out

    def set_rhs(self, b):
        lib().orc_gmg_set_rhs(self.h, dptr(np.ascontiguousarray(b, dtype=np.float64)))

    def set_u(self, u):
        lib().orc_gmg_set_u(self.h, dptr(np.ascontiguousarray(u, dtype=np.float64)))

    def u(self):
        out = self._arr()
        lib().orc_gmg_get_u(self.h, dptr(out))
        return out

    def apply(self, x):
        out = self._arr()
        lib().orc_gmg_apply(self.h, dptr(np.ascontiguousarray(x, dtype=np.float64)), dptr(out))
        return out

    def solve(self, tol=1e-10, maxit=100):
        it = lib().orc_gmg_solve(self.h, tol, maxit)
        hist = np.zeros(maxit + 2)
        n = lib().orc_gmg_history(self.h, dptr(hist), len(hist))
        return it, hist[:n]

    def restrict(self, level, fine):
        n = fine.shape[0] // 2
        out = self._arr(n)
        lib().orc_gmg_restrict(self.h, level, dptr(np.ascontiguousarray(fine)), dptr(out))
        return out

    def prolong_add(self, level, coarse, fine):
        f = np.ascontiguousarray(fine, dtype=np.float64).copy()
        lib().orc_gmg_prolong_add(self.h, level, dptr(np.ascontiguousarray(coarse)), dptr(f))
        return f
