"""Burgers benchmark driver over the parthenon_amd task API.

Mirrors BurgersDriver::MakeTaskCollection (benchmarks/burgers/burgers_driver.cpp:53-148),
including its flux-correction tasks (lines 89-96), with prolongation after the sets in the
library exchange order (AddBoundaryExchangeTasks, src/bvals/comms/boundary_communication.cpp:
406-445; the shipped benchmark never prolongates inside a stage) and
the low-storage RK2 integrator (src/time_integration/low_storage_integrator.cpp:67-83, stage
names "base", "1", "base") and EvolutionDriver's time-step control
(src/driver/driver.cpp:210-269): the first dt is the mesh estimate, later ones are
min(2 * dt_prev, estimate), within the <parthenon/time> limits and cut to end on tlim.
"""
import math
import sys

from . import (ApplyBoundaryConditionsOnCoarseOrFineMD, BoundaryType,
               LoadAndSendFluxCorrections, Mesh, Metadata, ProlongateBounds, Refinement,
               ReceiveBoundBufs, ReceiveFluxCorrections, SendBoundBufs, SetBounds,
               SetFluxCorrections, StartReceiveBoundBufs, StartReceiveFluxCorrections,
               TaskStatus, Update, burgers_package)

RK2_BETA = (1.0, 0.5)


class TimeLimits:
    """<parthenon/time> tlim, nlim and dt bounds with EvolutionDriver's defaults
    (src/driver/driver.hpp:66-94)."""

    def __init__(self, tlim=math.inf, nlim=-1, dt_init=sys.float_info.max, dt_init_force=False,
                 dt_user=sys.float_info.max, dt_force=-sys.float_info.max,
                 dt_floor=sys.float_info.min, dt_ceil=sys.float_info.max):
        self.tlim, self.nlim = tlim, nlim
        self.dt_init, self.dt_init_force = dt_init, dt_init_force
        self.dt_user, self.dt_force = dt_user, dt_force
        self.dt_floor, self.dt_ceil = dt_floor, dt_ceil


def set_global_time_step(dt, block_dt, time, ncycle, lim):
    """EvolutionDriver::SetGlobalTimeStep (driver.cpp:210-269, dt_factor 2): the new dt."""
    if lim.dt_force > 0.0:
        dt = lim.dt_force
    elif ncycle == 0 and lim.dt_init_force and lim.dt_init > 0.0:
        dt = lim.dt_init
    else:
        if dt < 0.1 * sys.float_info.max:
            dt *= 2.0
        if ncycle == 0:
            dt = min(dt, lim.dt_init)
        dt = min(dt, block_dt)
        dt = min(dt, lim.dt_user)
        dt = max(lim.dt_floor, min(dt, lim.dt_ceil))
    if time < lim.tlim and (lim.tlim - time) < dt:
        dt = lim.tlim - time
    return dt


def make_burgers_mesh(nx=(64, 64, 64), block=(16, 16, 16), nghost=4, num_scalars=4,
                      refine_spheres=(), refine_boxes=(), bcs=None, adaptive=None):
    """Mesh with the fields burgers_package::Initialize registers
    (benchmarks/burgers/burgers_package.cpp:76-94): U with FillGhost/Independent/WithFluxes and
    the OneCopy derived field. bcs: optional face conditions (inner x1, outer x1, ...,
    outer x3); a dimension whose faces are not periodic is not periodic. adaptive: optional
    dict(numlevel=..., derefine_count=..., criteria=[dict(field=, comp=, method=, refine_tol=,
    derefine_tol=, max_level=)]) for refinement = adaptive; the VIBE input
    (benchmarks/burgers/burgers.pin:21-55) is numlevel 2, derefine_count 10 and one
    derivative_order_1 criterion on U component 3 with tolerances 0.5 / 0.2."""
    periodic = (True, True, True)
    if bcs is not None:
        periodic = tuple(bcs[2 * d] == "periodic" and bcs[2 * d + 1] == "periodic"
                         for d in range(3))
    m = Mesh(nx, block, nghost, periodic=periodic, refine_spheres=refine_spheres,
             refine_boxes=refine_boxes, bcs=bcs)
    m.add_field("U", 3 + num_scalars,
                Metadata.FillGhost | Metadata.Independent | Metadata.WithFluxes)
    m.add_field("derived", 1, Metadata.OneCopy)
    if adaptive is not None:
        m.set_adaptive(adaptive.get("numlevel", 1), adaptive.get("derefine_count", 10))
        for c in adaptive.get("criteria", ()):
            m.add_refinement_criterion(**c)
    m.finalize()
    return m


class DeviceBurgersDriver:
    """The same cycle with every host decision hoisted out: one reconstruction launch, one fused
    flux/divergence/update launch and the exchange per stage, dt and time in device memory, and
    (use_graph) one captured hipGraph replayed per cycle."""

    def __init__(self, mesh, cfl=0.4, recon=burgers_package.WENO5, use_graph=True,
                 overlap=None, sweep=None, bit_exact=False, split=None):
        import ctypes as C
        from . import _check, lib
        self._C, self._check, self._lib = C, _check, lib
        self.mesh = mesh
        self.cfl = cfl
        self.recon = recon
        self.use_graph = use_graph
        self.base = mesh.mesh_data("base")
        self.one = mesh.mesh_data("1")
        # stage form: sweep = one block-sweep launch per stage (FMA-contracted, equal to the
        # others to rounding); otherwise the plane reconstruction (bit-exact), with overlap =
        # the stage input's exchange beside the ghost-free reconstruction on a second stream.
        # Left as None, initialize() times the forms and keeps the fastest; bit_exact limits
        # that choice to the bit-exact plane forms.
        if bit_exact:
            _check(lib().pth_burgers_set_bit_exact(mesh._h, 1), "set_bit_exact")
        if overlap is not None:
            _check(lib().pth_burgers_set_overlap(mesh._h, int(overlap)), "set_overlap")
        if sweep is not None:
            _check(lib().pth_burgers_set_sweep(mesh._h, int(sweep)), "set_sweep")
        if split is not None:  # the split sweep on refined 16^3 meshes (on by default)
            _check(lib().pth_burgers_set_sweep_split(mesh._h, int(split)), "set_sweep_split")

    def initialize(self, pgen="sinusoid", seed=0):
        if pgen == "sinusoid":
            burgers_package.ProblemGeneratorSinusoid(self.base, seed)
        elif pgen == "vibe":
            burgers_package.ProblemGenerator(self.base)
        self._check(self._lib().pth_burgers_init(self.mesh._h, self.cfl, self.recon), "init")

    def cycles(self, n):
        self._check(self._lib().pth_burgers_run_cycles(self.mesh._h, n, int(self.use_graph)),
                    "run_cycles")

    def synchronize(self):
        self._check(self._lib().pth_burgers_synchronize(self.mesh._h), "synchronize")

    def adaptive_cycles(self, n):
        """n cycles of EvolutionDriver::Execute on an adaptive mesh (driver.cpp:110-131): a
        cycle, Refinement::Tag of "base", the remesh and, when the mesh changed, FillDerived and
        the time step from the new blocks. Returns the number of remeshes."""
        changed = 0
        for _ in range(n):
            self.cycles(1)
            Refinement.Tag(self.base)
            if self.mesh.adapt():
                self._check(self._lib().pth_burgers_after_remesh(self.mesh._h), "after_remesh")
                changed += 1
        return changed

    def form(self):
        """The stage form in use: "sweep", "overlap" or "serial"."""
        C = self._C
        sw, ov = C.c_int(), C.c_int()
        self._lib().pth_burgers_form(self.mesh._h, C.byref(sw), C.byref(ov))
        return "sweep" if sw.value else ("overlap" if ov.value else "serial")

    def tuning(self):
        """Planes per workgroup of the three plane-reconstruction forms and the block sweep's
        waves per SIMD (picked at initialize)."""
        C = self._C
        out = (C.c_int * 4)()
        self._lib().pth_burgers_tuning(self.mesh._h, out)
        return list(out)

    def set_tuning(self, planes):
        C = self._C
        self._check(self._lib().pth_burgers_set_tuning(self.mesh._h, (C.c_int * 4)(*planes)),
                    "set_tuning")

    def clock(self):
        C = self._C
        dt, t = C.c_double(), C.c_double()
        self._check(self._lib().pth_burgers_clock(self.mesh._h, C.byref(dt), C.byref(t)), "clock")
        return dt.value, t.value

    def set_time_limits(self, tlim=math.inf, dt_init=sys.float_info.max, dt_init_force=False,
                        dt_user=sys.float_info.max, dt_force=-sys.float_info.max,
                        dt_floor=sys.float_info.min, dt_ceil=sys.float_info.max):
        """<parthenon/time> tlim and dt bounds (EvolutionDriver, src/driver/driver.hpp:66-94).
        Cycles queued after time reaches tlim leave the state and the clock unchanged."""
        C = self._C
        v = (C.c_double * 6)(tlim, dt_init, dt_user, dt_force, dt_floor, dt_ceil)
        self._check(self._lib().pth_burgers_set_time_limits(self.mesh._h, v, int(dt_init_force)),
                    "set_time_limits")


class BurgersDriver:
    def __init__(self, mesh, cfl=0.4, recon=burgers_package.WENO5, limits=None):
        self.mesh = mesh
        self.cfl = cfl
        self.recon = recon
        self.limits = limits or TimeLimits()
        self.base = mesh.mesh_data("base")
        self.one = mesh.mesh_data("1")
        self.dudt = mesh.mesh_data("dUdt")
        self.dt = sys.float_info.max  # <parthenon/time> dt default
        self.time = 0.0
        self.ncycle = 0

    def keep_going(self):
        """SimTime::KeepGoing (src/basic_types.hpp:241)."""
        lim = self.limits
        return self.time < lim.tlim and (lim.nlim < 0 or self.ncycle < lim.nlim)

    def _exchange(self, md):
        any_ = BoundaryType.any
        SendBoundBufs(md, BoundaryType.nonlocal_)
        SendBoundBufs(md, BoundaryType.local)
        if ReceiveBoundBufs(md, BoundaryType.local) != TaskStatus.complete:
            raise RuntimeError("local boundary buffers were not received")
        SetBounds(md, BoundaryType.local)
        ReceiveBoundBufs(md, BoundaryType.nonlocal_)
        SetBounds(md, BoundaryType.nonlocal_)
        # AddBoundaryExchangeTasks order (boundary_communication.cpp:432-445); the Burgers task
        # list applies the conditions to "sc1" after its exchange (burgers_driver.cpp:130-137)
        if self.mesh.multilevel:
            ApplyBoundaryConditionsOnCoarseOrFineMD(md, True)
            ProlongateBounds(md, any_)
        ApplyBoundaryConditionsOnCoarseOrFineMD(md, False)

    def initialize(self, pgen="sinusoid", seed=0):
        """Problem generator, boundary exchange, FillDerived, first time step
        (src/mesh/mesh.cpp:745-800, src/driver/driver.cpp:67-70)."""
        if pgen == "sinusoid":
            burgers_package.ProblemGeneratorSinusoid(self.base, seed)
        elif pgen == "vibe":
            burgers_package.ProblemGenerator(self.base)
        StartReceiveBoundBufs(self.base, BoundaryType.any)
        self._exchange(self.base)
        burgers_package.CalculateDerived(self.base)
        self.dt = set_global_time_step(self.dt, burgers_package.EstimateTimestepMesh(
            self.base, self.cfl), self.time, self.ncycle, self.limits)

    def stage(self, stage):
        beta = RK2_BETA[stage - 1]
        mc0 = self.base if stage == 1 else self.one
        mc1 = self.one if stage == 1 else self.base
        StartReceiveBoundBufs(mc1, BoundaryType.any)
        StartReceiveFluxCorrections(mc0)
        burgers_package.CalculateFluxes(mc0, self.recon)
        LoadAndSendFluxCorrections(mc0)
        if ReceiveFluxCorrections(mc0) != TaskStatus.complete:
            raise RuntimeError("flux corrections were not received")
        SetFluxCorrections(mc0)
        Update.FluxDivergence(mc0, self.dudt)
        Update.AverageIndependentData(mc0, self.base, beta)
        Update.UpdateIndependentData(mc0, self.dudt, beta * self.dt, mc1)
        self._exchange(mc1)
        burgers_package.CalculateDerived(mc1)
        if stage == 2:
            return burgers_package.EstimateTimestepMesh(mc1, self.cfl)
        return None

    def cycle(self):
        self.stage(1)
        new_dt = self.stage(2)
        self.ncycle += 1
        self.time += self.dt
        self.dt = set_global_time_step(self.dt, new_dt, self.time, self.ncycle, self.limits)

    def run(self, n):
        """EvolutionDriver::Execute's loop: at most n cycles while keep_going(); returns the
        number run."""
        q = 0
        while q < n and self.keep_going():
            self.cycle()
            q += 1
        return q

    def adaptive_cycle(self):
        """One cycle, then Refinement::Tag (burgers_driver.cpp:139-145) and
        LoadBalancingAndAdaptiveMeshRefinement; after a remesh FillDerived and
        InitializeBlockTimeSteps + SetGlobalTimeStep (driver.cpp:128-131). True if remeshed."""
        dt_used = self.dt
        self.cycle()
        Refinement.Tag(self.base)
        if not self.mesh.adapt():
            return False
        burgers_package.CalculateDerived(self.base)
        self.dt = set_global_time_step(dt_used, burgers_package.EstimateTimestepMesh(
            self.base, self.cfl), self.time, self.ncycle, self.limits)
        return True
