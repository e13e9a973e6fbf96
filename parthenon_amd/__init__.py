"""parthenon_amd — MI355X-native per-stage meshblock data path behind Parthenon's task API.

Host-side mirror of the reference interface for this path, over the C ABI in
include/parthenon_amd.h (libparthenon_amd.so, built in-tree). Names follow the reference:

  Mesh / MeshData                         src/mesh/mesh.hpp, src/interface/mesh_data.hpp
  SendBoundBufs, StartReceiveBoundBufs,   src/bvals/comms/bvals_in_one.hpp:41-93
  ReceiveBoundBufs, SetBounds,
  ProlongateBounds
  burgers_package.CalculateFluxes, ...    benchmarks/burgers/burgers_package.hpp:33-43
  Update.FluxDivergence, ...              src/interface/update.hpp:59-300

Every task returns a TaskStatus like the reference; a failure raises RuntimeError carrying
the library's message. There is no CPU fallback: if the HIP library is missing, importing
the compute API fails loudly.
"""
import ctypes as C
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparthenon_amd.so")
_LIB = None


class TaskStatus(enum.IntEnum):  # src/basic_types.hpp:59
    complete = 0
    incomplete = 1
    iterate = 2
    fail = 3


class BoundaryType(enum.IntEnum):  # src/basic_types.hpp:156
    any = 0
    local = 1
    nonlocal_ = 2
    flxcor_send = 3
    flxcor_recv = 4


class PDOpt(enum.IntFlag):  # src/interface/sparse_pack_base.hpp:46
    WithFluxes = 1
    Coarse = 2
    Flatten = 4


class Metadata(enum.IntFlag):  # the flags of src/interface/metadata.hpp used on this path
    FillGhost = 1
    Independent = 2
    WithFluxes = 4
    OneCopy = 8
    Vector = 16


class BoundaryCondition(enum.IntEnum):  # <parthenon/mesh> ix1_bc ... ox3_bc
    periodic = 0
    outflow = 1
    reflect = 2
    fixed_face = 3  # FixedFace with value 0 (the poisson_gmg package's condition)


class PackMode(enum.IntEnum):
    persistent = 0
    per_buffer = 1


_D = C.c_double
_I = C.c_int
_I64 = C.c_int64
_P = C.c_void_p
_DP = C.POINTER(C.c_double)


class PackView(C.Structure):
    """pth_pack_view of include/parthenon_amd.h: device pointers and shape of one pack."""
    _fields_ = [("data", C.c_void_p), ("flux", C.c_void_p), ("bounds", C.c_void_p),
                ("coords", C.c_void_p), ("geom", C.c_void_p), ("nblocks", C.c_int),
                ("nvar", C.c_int), ("nfields", C.c_int), ("options", C.c_int),
                ("ent", C.c_int * 3), ("is_", C.c_int), ("ie", C.c_int), ("js", C.c_int),
                ("je", C.c_int), ("ks", C.c_int), ("ke", C.c_int)]


class PoissonEquationC(C.Structure):
    """pth_poisson_equation of include/parthenon_amd.h."""
    _fields_ = [("D", C.c_double), ("diagonal_alpha", C.c_double), ("D_field", C.c_char_p)]


class MgParamsC(C.Structure):
    """pth_mg_params of include/parthenon_amd.h."""
    _fields_ = [("max_iters", C.c_int), ("residual_tolerance", C.c_double),
                ("precision", C.c_int), ("inner_cycles", C.c_int), ("omega", C.c_double),
                ("pre_sweeps", C.c_int), ("post_sweeps", C.c_int), ("smoother", C.c_int)]


_SIGNATURES = {
    "pth_mesh_create": (_P, [C.POINTER(_I), C.POINTER(_I), _I, _DP, _DP, C.POINTER(_I)]),
    "pth_mesh_refine_box": (_I, [_P, _DP, _I]),
    "pth_mesh_set_boundary": (_I, [_P, _I, _I]),
    "pth_apply_boundary_conditions": (_I, [_P, _I]),
    "pth_mesh_refine_sphere": (_I, [_P, _DP, _D, _I]),
    "pth_mesh_add_field": (_I, [_P, C.c_char_p, _I, C.c_uint]),
    "pth_mesh_plan": (_I, [_P]),
    "pth_mesh_register_refinement_ops": (_I, [_P, C.c_char_p, _I, _I]),
    "pth_mesh_finalize": (_I, [_P]),
    "pth_mesh_destroy": (None, [_P]),
    "pth_mesh_set_stream": (_I, [_P, _P]),
    "pth_mesh_synchronize": (_I, [_P]),
    "pth_mesh_set_pack_mode": (_I, [_P, _I]),
    "pth_mesh_nblocks": (_I, [_P]),
    "pth_mesh_multilevel": (_I, [_P]),
    "pth_mesh_block_info": (_I, [_P, _I, C.POINTER(_I64), _DP]),
    "pth_mesh_index_shapes": (_I, [_P, C.POINTER(_I)]),
    "pth_mesh_neighbors": (_I, [_P, _I, C.POINTER(_I64), _I]),
    "pth_mesh_region_boxes": (_I, [_P, _I, C.POINTER(_I), _I]),
    "pth_mesh_exchange_counts": (_I, [_P, _I, C.POINTER(_I64), C.POINTER(_I64)]),
    "pth_meshdata_get": (_P, [_P, C.c_char_p]),
    "pth_field_pointers": (_I, [_P, C.c_char_p, C.POINTER(_DP), C.POINTER(_DP), C.POINTER(_DP)]),
    "pth_pack_get": (_P, [_P, C.POINTER(C.c_char_p), _I, C.c_uint]),
    "pth_pack_view_get": (_I, [_P, C.POINTER(PackView)]),
    "pth_pack_bounds": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "pth_field_upload": (_I, [_P, C.c_char_p, _I, _DP, C.c_size_t]),
    "pth_field_download": (_I, [_P, C.c_char_p, _I, _DP, C.c_size_t]),
    "pth_send_bound_bufs": (_I, [_P, _I]),
    "pth_start_receive_bound_bufs": (_I, [_P, _I]),
    "pth_receive_bound_bufs": (_I, [_P, _I]),
    "pth_set_bounds": (_I, [_P, _I]),
    "pth_prolongate_bounds": (_I, [_P, _I]),
    "pth_burgers_calculate_fluxes": (_I, [_P, _I]),
    "pth_flux_divergence": (_I, [_P, _P]),
    "pth_weighted_sum_independent": (_I, [_P, _P, _D, _D, _P]),
    "pth_burgers_calculate_derived": (_I, [_P]),
    "pth_burgers_estimate_timestep": (_I, [_P, _D, _DP]),
    "pth_burgers_estimate_timestep_noalloc": (_I, [_P, _D, _P]),
    "pth_burgers_pgen_sinusoid": (_I, [_P, C.c_uint64]),
    "pth_burgers_pgen_vibe": (_I, [_P]),
    "pth_burgers_mass_history": (_I, [_P, _DP]),
    "pth_exchange_fused": (_I, [_P, _I]),
    "pth_burgers_init": (_I, [_P, _D, _I]),
    "pth_burgers_run_cycles": (_I, [_P, _I, _I]),
    "pth_burgers_clock": (_I, [_P, _DP, _DP]),
    "pth_burgers_set_time_limits": (_I, [_P, _DP, _I]),
    "pth_burgers_set_overlap": (_I, [_P, _I]),
    "pth_burgers_set_sweep": (_I, [_P, _I]),
    "pth_burgers_set_sweep_split": (_I, [_P, _I]),
    "pth_burgers_set_bit_exact": (_I, [_P, _I]),
    "pth_burgers_form": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "pth_burgers_tuning": (_I, [_P, C.POINTER(_I)]),
    "pth_burgers_set_tuning": (_I, [_P, C.POINTER(_I)]),
    "pth_burgers_synchronize": (_I, [_P]),
    "pth_mesh_set_adaptive": (_I, [_P, _I, _I]),
    "pth_mesh_add_refinement_criterion": (_I, [_P, _I, C.c_char_p, _I, _D, _D, _I]),
    "pth_refinement_tag": (_I, [_P]),
    "pth_mesh_set_refinement": (_I, [_P, _I, _I]),
    "pth_mesh_refinement_flags": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "pth_mesh_root_level": (_I, [_P]),
    "pth_mesh_adapt": (_I, [_P, C.POINTER(_I)]),
    "pth_burgers_after_remesh": (_I, [_P]),
    "pth_gmg_create": (_P, [_I, _I, _D, _I, _I]),
    "pth_gmg_destroy": (None, [_P]),
    "pth_gmg_levels": (_I, [_P]),
    "pth_gmg_set_rhs_blobs": (_I, [_P, C.c_uint64]),
    "pth_gmg_set_rhs": (_I, [_P, _DP, C.c_size_t]),
    "pth_gmg_get_rhs": (_I, [_P, _DP, C.c_size_t]),
    "pth_gmg_get_solution": (_I, [_P, _DP, C.c_size_t]),
    "pth_gmg_reset": (_I, [_P]),
    "pth_gmg_solve": (_I, [_P, _D, _I, C.POINTER(_I), _DP]),
    "pth_gmg_set_inner_cycles": (_I, [_P, _I]),
    "pth_gmg_set_column_sweeps": (_I, [_P, _I]),
    "pth_gmg_server_set_column_sweeps": (_I, [_P, _I]),
    "pth_gmg_server_set_hbm_reserve": (_I, [_P, C.c_int64]),
    "pth_device_memory": (_I, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pth_gmg_server_set_inner_cycles": (_I, [_P, _I]),
    "pth_gmg_history": (_I, [_P, _DP, _I]),
    "pth_solver_copy_data": (_I, [_P, C.c_char_p, C.c_char_p]),
    "pth_solver_add_fields_and_store": (_I, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _D, _D]),
    "pth_solver_set_to_zero": (_I, [_P, C.c_char_p]),
    "pth_solver_dot_product": (_I, [_P, C.c_char_p, C.c_char_p, _DP]),
    "pth_poisson_ax": (_I, [_P, C.POINTER(PoissonEquationC), C.c_char_p, C.c_char_p]),
    "pth_mg_create": (_P, [_P, C.c_char_p, C.c_char_p, C.POINTER(MgParamsC),
                           C.POINTER(PoissonEquationC)]),
    "pth_mg_destroy": (None, [_P]),
    "pth_mg_setup": (_I, [_P, _P]),
    "pth_mg_solve": (_I, [_P, _P]),
    "pth_mg_result": (_I, [_P, C.POINTER(_I), _DP]),
    "pth_mg_history": (_I, [_P, _DP, _I]),
    "pth_mg_levels": (_I, [_P]),
    "pth_mg_set_graph": (_I, [_P, _I]),
    "pth_gmg_server_create": (_P, [_I, _D, _I, _I, _D, _I, _I64, _I, _I]),
    "pth_gmg_server_destroy": (None, [_P]),
    "pth_gmg_server_submit": (_I, [_P, _I, C.c_uint64, _D, C.POINTER(_I)]),
    "pth_gmg_server_run": (_I, [_P, _I]),
    "pth_gmg_server_result": (_I, [_P, _I, _DP, C.POINTER(_I)]),
    "pth_gmg_server_solution": (_I, [_P, _I, _DP, C.c_size_t]),
    "pth_gmg_server_stats": (_I, [_P, C.POINTER(_I64), _DP]),
    "pth_profile_enable": (_I, [_I]),
    "pth_profile_report": (_I, [C.c_char_p, _I]),
    "pth_last_error": (C.c_char_p, []),
    "pth_version": (C.c_char_p, []),
}


def lib():
    """Load libparthenon_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGNATURES)


def _check(status, what):
    if status == TaskStatus.fail:
        msg = lib().pth_last_error()
        raise RuntimeError(f"{what} failed: {msg.decode() if msg else 'unknown error'}")
    return TaskStatus(status)


def _dptr(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays crossing the C ABI must be contiguous float64")
    return a.ctypes.data_as(_DP)


class Mesh:
    """A uniform or statically refined Cartesian mesh of equal-size meshblocks."""

    def __init__(self, nx, block, nghost, xmin=(-0.5, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                 periodic=(True, True, True), refine_boxes=(), refine_spheres=(), bcs=None):
        """bcs: optional conditions of the 6 faces (inner x1, outer x1, ..., outer x3), each
        a BoundaryCondition or its name; faces of non-periodic dimensions default to
        outflow, faces of periodic dimensions must stay periodic."""
        L = lib()
        arr3 = lambda v, t: (t * 3)(*v)
        self._h = L.pth_mesh_create(arr3(nx, _I), arr3(block, _I), nghost, arr3(xmin, _D),
                                    arr3(xmax, _D), arr3([int(bool(p)) for p in periodic], _I))
        if not self._h:
            raise RuntimeError("pth_mesh_create failed: " + L.pth_last_error().decode())
        for face, bc in enumerate(bcs or ()):
            bc = BoundaryCondition[bc] if isinstance(bc, str) else BoundaryCondition(bc)
            if bc != BoundaryCondition.periodic:
                _check(L.pth_mesh_set_boundary(self._h, face, int(bc)), "set_boundary")
        for box, level in refine_boxes:
            _check(L.pth_mesh_refine_box(self._h, (_D * 6)(*box), level), "refine_box")
        for center, radius, level in refine_spheres:
            _check(L.pth_mesh_refine_sphere(self._h, arr3(center, _D), radius, level),
                   "refine_sphere")
        self.fields = {}
        self._data = {}
        self.finalized = False

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            self._data.clear()
            _LIB.pth_mesh_destroy(self._h)
            self._h = None

    # -- setup --
    def add_field(self, name, ncomp=1, flags=Metadata.FillGhost | Metadata.Independent):
        _check(lib().pth_mesh_add_field(self._h, name.encode(), ncomp, int(flags)), "add_field")
        self.fields[name] = (ncomp, int(flags))

    def register_refinement_ops(self, name, prolong="minmod"):
        """Metadata::RegisterRefinementOps for field name: shared prolongation "minmod"
        (ProlongateSharedMinMod, the default) or "linear" (ProlongateSharedLinear);
        restriction is RestrictAverage. Before finalize."""
        op = {"minmod": 0, "linear": 1}[prolong]
        _check(lib().pth_mesh_register_refinement_ops(self._h, name.encode(), op, 0),
               "register_refinement_ops")

    def plan(self):
        _check(lib().pth_mesh_plan(self._h), "plan")
        self._shapes()

    def finalize(self):
        _check(lib().pth_mesh_finalize(self._h), "finalize")
        self._shapes()
        self.finalized = True

    def _shapes(self):
        s = (_I * 12)()
        _check(lib().pth_mesh_index_shapes(self._h, s), "index_shapes")
        self.n, self.ng = tuple(s[0:3]), tuple(s[3:6])
        self.cn, self.cng = tuple(s[6:9]), tuple(s[9:12])
        self.ent = tuple(self.n[d] + 2 * self.ng[d] for d in range(3))
        self.cent = tuple(self.cn[d] + 2 * self.cng[d] for d in range(3))
        self.nblocks = lib().pth_mesh_nblocks(self._h)
        self.multilevel = bool(lib().pth_mesh_multilevel(self._h))

    # -- queries --
    def block(self, gid):
        out = (_I64 * 4)()
        x = (_D * 6)()
        _check(lib().pth_mesh_block_info(self._h, gid, out, x), "block_info")
        return dict(level=out[0], l=tuple(out[1:4]), xmin=tuple(x[0:3]), xmax=tuple(x[3:6]))

    def neighbors(self, gid):
        buf = (_I64 * (8 * 64))()
        n = lib().pth_mesh_neighbors(self._h, gid, buf, 64)
        return [dict(gid=buf[8 * q], level=buf[8 * q + 1], origin=tuple(buf[8 * q + 2:8 * q + 5]),
                     off=tuple(buf[8 * q + 5:8 * q + 8])) for q in range(n)]

    def region_boxes(self, gid):
        buf = (_I * (7 * 512))()
        n = lib().pth_mesh_region_boxes(self._h, gid, buf, 512)
        return np.array(buf[:7 * n], dtype=np.int64).reshape(n, 7)

    def exchange_counts(self, bt=BoundaryType.any):
        nb, ne = _I64(), _I64()
        _check(lib().pth_mesh_exchange_counts(self._h, int(bt), C.byref(nb), C.byref(ne)),
               "exchange_counts")
        return nb.value, ne.value

    def field_shape(self, name):
        ncomp = self.fields[name][0]
        return (self.nblocks, ncomp, self.ent[2], self.ent[1], self.ent[0])

    def coarse_shape(self, name):
        ncomp = self.fields[name][0]
        return (self.nblocks, ncomp, self.cent[2], self.cent[1], self.cent[0])

    # -- runtime --
    def mesh_data(self, name="base"):
        """MeshData<Real> for a named container (mesh_data.GetOrAdd)."""
        if name not in self._data:
            h = lib().pth_meshdata_get(self._h, name.encode())
            if not h:
                raise RuntimeError("pth_meshdata_get failed: " + lib().pth_last_error().decode())
            self._data[name] = MeshData(self, name, h)
        return self._data[name]

    def set_stream(self, stream_ptr):
        _check(lib().pth_mesh_set_stream(self._h, C.c_void_p(stream_ptr)), "set_stream")

    def set_pack_mode(self, mode):
        _check(lib().pth_mesh_set_pack_mode(self._h, int(mode)), "set_pack_mode")

    def synchronize(self):
        _check(lib().pth_mesh_synchronize(self._h), "synchronize")

    # -- adaptive refinement (<parthenon/mesh> refinement = adaptive) --
    def set_adaptive(self, numlevel, derefine_count=10):
        """numlevel and derefine_count of <parthenon/mesh>; before finalize."""
        _check(lib().pth_mesh_set_adaptive(self._h, numlevel, derefine_count), "set_adaptive")

    def add_refinement_criterion(self, field, comp=0, method="derivative_order_1",
                                 refine_tol=0.5, derefine_tol=0.05, max_level=0):
        """One <parthenon/refinementN> block (field, vector_i, method, tolerances)."""
        m = {"derivative_order_1": 1, "derivative_order_2": 2}.get(method, method)
        _check(lib().pth_mesh_add_refinement_criterion(self._h, int(m), field.encode(), comp,
                                                       refine_tol, derefine_tol, max_level),
               "add_refinement_criterion")

    @property
    def root_level(self):
        return lib().pth_mesh_root_level(self._h)

    def set_refinement(self, gid, tag):
        """MeshRefinement::SetRefinement(AmrTag) of block gid (-1, 0, 1)."""
        _check(lib().pth_mesh_set_refinement(self._h, gid, int(tag)), "set_refinement")

    def refinement_flags(self):
        """(refine_flag_, deref_count_) of every block."""
        f, c = (_I * self.nblocks)(), (_I * self.nblocks)()
        _check(lib().pth_mesh_refinement_flags(self._h, f, c), "refinement_flags")
        return np.array(f[:], dtype=np.int32), np.array(c[:], dtype=np.int32)

    def adapt(self):
        """Mesh::LoadBalancingAndAdaptiveMeshRefinement: True when the mesh changed. Block
        ids and counts change; MeshData objects stay valid."""
        mod = _I()
        _check(lib().pth_mesh_adapt(self._h, C.byref(mod)), "adapt")
        if mod.value:
            self._shapes()
        return bool(mod.value)

    def locs(self):
        """(level, (lx1, lx2, lx3)) of every block in gid order."""
        return [(b["level"], b["l"]) for b in map(self.block, range(self.nblocks))]


class Refinement:  # src/amr_criteria/refinement_package.hpp
    @staticmethod
    def Tag(md):
        return _check(lib().pth_refinement_tag(md._h), "Refinement::Tag")


def LoadBalancingAndAdaptiveMeshRefinement(mesh):
    """src/mesh/mesh-amr_loadbalance.cpp:324-351 on one GPU; returns Mesh::modified."""
    return mesh.adapt()


class MeshData:
    def __init__(self, mesh, name, handle):
        self.mesh = mesh
        self.name = name
        self._h = handle

    def _shape(self, field, which):
        if which == 1:
            return self.mesh.coarse_shape(field)
        s = self.mesh.field_shape(field)
        return (3,) + s if which == 2 else s

    def upload(self, field, array, which=0):
        a = np.ascontiguousarray(array, dtype=np.float64)
        if a.shape != self._shape(field, which):
            raise ValueError(f"shape {a.shape} != {self._shape(field, which)}")
        _check(lib().pth_field_upload(self._h, field.encode(), which, _dptr(a), a.size), "upload")

    def download(self, field, which=0):
        out = np.empty(self._shape(field, which), dtype=np.float64)
        _check(lib().pth_field_download(self._h, field.encode(), which, _dptr(out), out.size),
               "download")
        return out

    def pack(self, fields, options=0):
        """MakePackDescriptor(md, fields, {}, options).GetPack(md) / PackVariables: the pack's
        handle and its device view (PackView); options from PDOpt."""
        names = (C.c_char_p * len(fields))(*[f.encode() for f in fields])
        h = lib().pth_pack_get(self._h, names, len(fields), int(options))
        if not h:
            raise RuntimeError(f"pack failed: {lib().pth_last_error().decode()}")
        view = PackView()
        _check(lib().pth_pack_view_get(h, C.byref(view)), "pack view")
        return h, view

    def pack_bounds(self, handle, field):
        lo, hi = _I(), _I()
        _check(lib().pth_pack_bounds(handle, field, C.byref(lo), C.byref(hi)), "pack bounds")
        return lo.value, hi.value

    def pointers(self, field):
        d, c, f = _DP(), _DP(), _DP()
        _check(lib().pth_field_pointers(self._h, field.encode(), C.byref(d), C.byref(c),
                                        C.byref(f)), "pointers")
        addr = lambda p: C.cast(p, C.c_void_p).value
        return addr(d), addr(c), addr(f)


# ---- boundary communication tasks (src/bvals/comms/bvals_in_one.hpp) ----
def SendBoundBufs(md, bt=BoundaryType.any):
    return _check(lib().pth_send_bound_bufs(md._h, int(bt)), "SendBoundBufs")


def StartReceiveBoundBufs(md, bt=BoundaryType.any):
    return _check(lib().pth_start_receive_bound_bufs(md._h, int(bt)), "StartReceiveBoundBufs")


def ReceiveBoundBufs(md, bt=BoundaryType.any):
    return _check(lib().pth_receive_bound_bufs(md._h, int(bt)), "ReceiveBoundBufs")


def SetBounds(md, bt=BoundaryType.any):
    return _check(lib().pth_set_bounds(md._h, int(bt)), "SetBounds")


def ProlongateBounds(md, bt=BoundaryType.any):
    return _check(lib().pth_prolongate_bounds(md._h, int(bt)), "ProlongateBounds")


# physical boundary conditions (src/bvals/boundary_conditions.hpp:40-50)
def ApplyBoundaryConditionsOnCoarseOrFineMD(md, coarse):
    return _check(lib().pth_apply_boundary_conditions(md._h, int(bool(coarse))),
                  "ApplyBoundaryConditionsOnCoarseOrFineMD")


def ApplyBoundaryConditionsMD(md):
    return ApplyBoundaryConditionsOnCoarseOrFineMD(md, False)


# flux-correction tasks (src/bvals/comms/bvals_in_one.hpp:69-80)
def StartReceiveFluxCorrections(md):
    return StartReceiveBoundBufs(md, BoundaryType.flxcor_recv)


def LoadAndSendFluxCorrections(md):
    return SendBoundBufs(md, BoundaryType.flxcor_send)


def ReceiveFluxCorrections(md):
    return ReceiveBoundBufs(md, BoundaryType.flxcor_recv)


def SetFluxCorrections(md):
    return SetBounds(md, BoundaryType.flxcor_recv)


class burgers_package:  # benchmarks/burgers/burgers_package.hpp
    WENO5 = 0
    Linear = 1

    @staticmethod
    def CalculateFluxes(md, recon=0):
        return _check(lib().pth_burgers_calculate_fluxes(md._h, recon), "CalculateFluxes")

    @staticmethod
    def CalculateDerived(md):
        return _check(lib().pth_burgers_calculate_derived(md._h), "CalculateDerived")

    @staticmethod
    def EstimateTimestepMesh(md, cfl):
        dt = _D()
        _check(lib().pth_burgers_estimate_timestep(md._h, cfl, C.byref(dt)), "EstimateTimestep")
        return dt.value

    @staticmethod
    def ProblemGeneratorSinusoid(md, seed):
        return _check(lib().pth_burgers_pgen_sinusoid(md._h, seed), "pgen_sinusoid")

    @staticmethod
    def ProblemGenerator(md):
        return _check(lib().pth_burgers_pgen_vibe(md._h), "pgen_vibe")

    @staticmethod
    def MassHistory(md):
        """The 8 octant masses (history variables "MS Mass 0..7")."""
        out = (_D * 8)()
        _check(lib().pth_burgers_mass_history(md._h, out), "MassHistory")
        return list(out)


class Update:  # src/interface/update.hpp
    @staticmethod
    def FluxDivergence(md_in, md_dudt):
        return _check(lib().pth_flux_divergence(md_in._h, md_dudt._h), "FluxDivergence")

    @staticmethod
    def WeightedSumData(in1, in2, w1, w2, out):
        return _check(lib().pth_weighted_sum_independent(in1._h, in2._h, w1, w2, out._h),
                      "WeightedSumData")

    @staticmethod
    def AverageIndependentData(c1, c2, wgt1):
        return Update.WeightedSumData(c1, c2, wgt1, 1.0 - wgt1, c1)

    @staticmethod
    def UpdateIndependentData(md_in, dudt, dt, out):
        return Update.WeightedSumData(md_in, dudt, 1.0, dt, out)


def device_memory():
    """(free, total) bytes of the current device's memory."""
    f, t = C.c_int64(), C.c_int64()
    _check(lib().pth_device_memory(C.byref(f), C.byref(t)), "device_memory")
    return f.value, t.value


def profile_enable(on=True):
    """Bracket every kernel launch with HIP events on its stream."""
    _check(lib().pth_profile_enable(1 if on else 0), "profile_enable")


def profile_report():
    """{kernel name: (launches, total_ms)} for the launches recorded since the last report."""
    buf = C.create_string_buffer(1 << 16)
    lib().pth_profile_report(buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, count, ms = line.split()
        out[name] = (int(count), float(ms))
    return out


def exchange_fused(md, bt=BoundaryType.any):
    """Send+Receive+Set(+coarse conditions, Prolongate), conditions in one call with the buffer
    stage skipped (one GPU)."""
    return _check(lib().pth_exchange_fused(md._h, int(bt)), "exchange_fused")


def exchange(md, bt=BoundaryType.any):
    """Send, receive, set (+ coarse-buffer conditions and prolongation on multilevel meshes),
    boundary conditions: the library task order of AddBoundaryExchangeTasks
    (src/bvals/comms/boundary_communication.cpp:406-445)."""
    SendBoundBufs(md, bt)
    if ReceiveBoundBufs(md, bt) != TaskStatus.complete:
        raise RuntimeError("ReceiveBoundBufs did not complete after SendBoundBufs")
    SetBounds(md, bt)
    if md.mesh.multilevel:
        ApplyBoundaryConditionsOnCoarseOrFineMD(md, True)
        ProlongateBounds(md, bt)
    ApplyBoundaryConditionsOnCoarseOrFineMD(md, False)
