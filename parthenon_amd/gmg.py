"""Geometric-multigrid Poisson solver and its serving loop, over the C ABI.

`MGSolver` mirrors the reference's solvers::MGSolver<u, rhs, PoissonEquation>
(src/solvers/mg_solver.hpp:37-521, MGParams :37-58) for the uniform N^3 grid of
example/poisson_gmg on [0, 1]^3 with FixedFace-0 boundaries. precision="mixed" runs the
V-cycles in FP32 inside an FP64 iterative-refinement loop; precision="fp64" is the same solver
in FP64. `Server` admits independent solves as they arrive and batches their V-cycle stages.
"""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _check, _dptr, lib

_PREC = {"mixed": 0, "fp64": 1}


@dataclass
class MGParams:  # src/solvers/mg_solver.hpp:37-58, with the smoother weight explicit
    max_iters: int = 100
    residual_tolerance: float = 1e-10   # relative: ||b - A u|| / ||b||
    omega: float = 2.0 / 3.0
    pre_sweeps: int = 2
    post_sweeps: int = 2
    inner_cycles: int = 1               # V-cycles per FP64 refinement step
    column_sweeps: bool = True          # Jacobi by z columns where the level size allows


class MGSolver:
    def __init__(self, n, precision="mixed", params=None):
        self.n = n
        self.params = params or MGParams()
        p = self.params
        h = lib().pth_gmg_create(n, _PREC[precision], p.omega, p.pre_sweeps, p.post_sweeps)
        if not h:
            raise RuntimeError("pth_gmg_create: " + lib().pth_last_error().decode())
        self._h = h
        self.levels = lib().pth_gmg_levels(h)
        _check(lib().pth_gmg_set_inner_cycles(h, p.inner_cycles), "set_inner_cycles")
        _check(lib().pth_gmg_set_column_sweeps(h, int(p.column_sweeps)), "set_column_sweeps")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pth_gmg_destroy(self._h)
            self._h = None

    def set_rhs_blobs(self, seed):
        _check(lib().pth_gmg_set_rhs_blobs(self._h, seed), "set_rhs_blobs")

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().pth_gmg_set_rhs(self._h, _dptr(b), b.size), "set_rhs")

    def rhs(self):
        out = np.empty((self.n,) * 3)
        _check(lib().pth_gmg_get_rhs(self._h, _dptr(out), out.size), "get_rhs")
        return out

    def solution(self):
        out = np.empty((self.n,) * 3)
        _check(lib().pth_gmg_get_solution(self._h, _dptr(out), out.size), "get_solution")
        return out

    def reset(self):
        _check(lib().pth_gmg_reset(self._h), "reset")

    def solve(self):
        it, rel = C.c_int(), C.c_double()
        _check(lib().pth_gmg_solve(self._h, self.params.residual_tolerance,
                                   self.params.max_iters, C.byref(it), C.byref(rel)), "solve")
        return it.value, rel.value

    def history(self):
        buf = (C.c_double * 1024)()
        n = lib().pth_gmg_history(self._h, buf, 1024)
        return np.array(buf[:min(n, 1024)])


class Server:
    """Continuous-batching request loop: admits solves by free device memory (and an optional
    cap on the bytes held, budget_bytes > 0) and free slots (max_active)."""

    def __init__(self, precision="mixed", params=None, budget_bytes=0, max_active=64,
                 keep_solutions=False, hbm_reserve=None):
        p = params or MGParams()
        self.params = p
        h = lib().pth_gmg_server_create(_PREC[precision], p.omega, p.pre_sweeps, p.post_sweeps,
                                        p.residual_tolerance, p.max_iters, budget_bytes,
                                        max_active, int(keep_solutions))
        if not h:
            raise RuntimeError("pth_gmg_server_create: " + lib().pth_last_error().decode())
        self._h = h
        self.sizes = {}
        _check(lib().pth_gmg_server_set_inner_cycles(h, p.inner_cycles), "set_inner_cycles")
        _check(lib().pth_gmg_server_set_column_sweeps(h, int(p.column_sweeps)),
               "set_column_sweeps")
        if hbm_reserve is not None:
            _check(lib().pth_gmg_server_set_hbm_reserve(h, int(hbm_reserve)), "set_hbm_reserve")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pth_gmg_server_destroy(self._h)
            self._h = None

    def submit(self, n, seed, arrival_s=0.0):
        rid = C.c_int()
        _check(lib().pth_gmg_server_submit(self._h, n, seed, arrival_s, C.byref(rid)), "submit")
        self.sizes[rid.value] = n
        return rid.value

    def run(self, realtime=True):
        _check(lib().pth_gmg_server_run(self._h, int(realtime)), "run")

    def result(self, rid):
        out = (C.c_double * 5)()
        it = C.c_int()
        _check(lib().pth_gmg_server_result(self._h, rid, out, C.byref(it)), "result")
        return dict(arrival=out[0], admit=out[1], done=out[2], latency=out[3],
                    rel_residual=out[4], iters=it.value)

    def solution(self, rid):
        n = self.sizes[rid]
        out = np.empty((n, n, n))
        _check(lib().pth_gmg_server_solution(self._h, rid, _dptr(out), out.size), "solution")
        return out

    def stats(self):
        out = (C.c_int64 * 3)()
        busy = C.c_double()
        _check(lib().pth_gmg_server_stats(self._h, out, C.byref(busy)), "stats")
        return dict(iterations=out[0], max_batch=out[1], bytes_held=out[2], busy_s=busy.value)


# ---- solver utilities, the Poisson equation and MGSolver over MeshData ----------------------
def copy_data(md, src, dst):
    """solvers::utils::CopyData<src, dst> (src/solvers/solver_utils.hpp:151-178)."""
    _check(lib().pth_solver_copy_data(md._h, src.encode(), dst.encode()), "CopyData")


def add_fields_and_store(md, a, b, out, wa=1.0, wb=1.0):
    """solvers::utils::AddFieldsAndStore<a, b, out> (solver_utils.hpp:180-222)."""
    _check(lib().pth_solver_add_fields_and_store(md._h, a.encode(), b.encode(), out.encode(),
                                                 wa, wb), "AddFieldsAndStore")


def set_to_zero(md, var):
    """solvers::utils::SetToZero<var> (solver_utils.hpp:224-250): interior cells."""
    _check(lib().pth_solver_set_to_zero(md._h, var.encode()), "SetToZero")


def dot_product(md, a, b):
    """solvers::utils::DotProduct<a, b> (solver_utils.hpp:252-301) on one GPU."""
    out = C.c_double()
    _check(lib().pth_solver_dot_product(md._h, a.encode(), b.encode(), C.byref(out)),
           "DotProduct")
    return out.value


@dataclass
class PoissonEquation:  # example/poisson_gmg/poisson_equation.hpp:41-205
    D: float = 1.0                 # constant D, or
    diagonal_alpha: float = 0.0
    D_field: Optional[str] = None  # a 3-component field: component d = D on the lower face
                                   # along d (the reference's face field D)

    def _c(self):
        from . import PoissonEquationC
        return PoissonEquationC(self.D, self.diagonal_alpha,
                                self.D_field.encode() if self.D_field else None)

    def Ax(self, md, x, out):
        """PoissonEquation::Ax<x, out> (:41-53): needs x's ghosts filled."""
        e = self._c()
        _check(lib().pth_poisson_ax(md._h, C.byref(e), x.encode(), out.encode()), "Ax")


_SMOOTHER = {"jacobi": 0, "SRJ1": 1, "SRJ2": 2, "SRJ3": 3}


@dataclass
class MeshMGParams:  # src/solvers/mg_solver.hpp:37-58, with the smoother explicit
    max_iters: int = 1000
    residual_tolerance: float = 1e-12   # RMS residual sqrt(sum r^2 / ncells), as the reference
    precision: str = "mixed"            # "fp64" on meshes other than the uniform N^3 cube
    inner_cycles: int = 2               # FP32 V-cycles per FP64 refinement step (mixed)
    omega: float = 2.0 / 3.0            # weighted Jacobi ("jacobi")
    pre_sweeps: int = 2
    post_sweeps: int = 2
    smoother: str = "jacobi"            # or the reference's "SRJ1" / "SRJ2" / "SRJ3"


class MeshMGSolver:
    """MGSolver<u, rhs, PoissonEquation> over the blocks of a MeshData (mg_solver.hpp:37-520):
    setup() is AddSetupTasks, solve() is AddTasks run to completion. A uniform 3-D N^3 cube
    with fixed_face faces runs on whole-level arrays (mixed precision or FP64); any other mesh
    (static refinement, 1-D / 2-D, periodic faces) on a hierarchy of block meshes in FP64."""

    def __init__(self, mesh, u, rhs, params=None, equation=None):
        from . import MgParamsC
        self.params = params or MeshMGParams()
        self.equation = equation or PoissonEquation()
        p = self.params
        cp = MgParamsC(p.max_iters, p.residual_tolerance, _PREC[p.precision], p.inner_cycles,
                       p.omega, p.pre_sweeps, p.post_sweeps, _SMOOTHER[p.smoother])
        e = self.equation._c()
        h = lib().pth_mg_create(mesh._h, u.encode(), rhs.encode(), C.byref(cp), C.byref(e))
        if not h:
            raise RuntimeError("pth_mg_create: " + lib().pth_last_error().decode())
        self._h = h
        self._mesh = mesh  # the solver refers to the mesh: keep it alive

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pth_mg_destroy(self._h)
            self._h = None

    def setup(self, md):
        _check(lib().pth_mg_setup(self._h, md._h), "MGSolver::AddSetupTasks")

    def solve(self, md):
        """Returns (V-cycles, final RMS residual)."""
        _check(lib().pth_mg_solve(self._h, md._h), "MGSolver::AddTasks")
        return self.result()

    def result(self):
        it, rms = C.c_int(), C.c_double()
        _check(lib().pth_mg_result(self._h, C.byref(it), C.byref(rms)), "MGSolver result")
        return it.value, rms.value

    def history(self):
        buf = (C.c_double * 1024)()
        n = lib().pth_mg_history(self._h, buf, 1024)
        return np.array(buf[:min(n, 1024)])

    def levels(self):
        """Levels of the hierarchy, finest first."""
        return lib().pth_mg_levels(self._h)

    def set_graph(self, on):
        """Block-mesh hierarchy: replay V-cycles from one captured hipGraph (on, the default)
        or launch every V-cycle eagerly (off); same results bit for bit."""
        _check(lib().pth_mg_set_graph(self._h, int(on)), "set_graph")
