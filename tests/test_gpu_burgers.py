"""GPU parity of the Burgers benchmark stencils and full RK2 cycles through the C ABI, against
the oracle on the same seeded inputs. Tolerance from the path's contract: FP64 flux updates
within 1e-13 relative; ghost copies are exact, and so are dt and time for the task and
plane-tiled forms (the block sweep is built with FMA contraction: within 1e-13)."""
import numpy as np
import pytest

import parthenon_amd as pa
from parthenon_amd.burgers import BurgersDriver, make_burgers_mesh
from oracle_lib import OracleBurgers, OracleMesh

pytestmark = pytest.mark.gpu

RTOL = 1e-13

# stage forms of the device driver: one block-sweep launch; plane-tiled reconstruction with
# the exchange on a second stream; plane-tiled reconstruction then update then exchange
FORMS = {"sweep": dict(sweep=True), "overlap": dict(sweep=False, overlap=True),
         "serial": dict(sweep=False, overlap=False)}


def pair(nx=(32, 32, 32), block=(16, 16, 16), g=4, ns=4, cfl=0.4, recon=0, **kw):
    m = make_burgers_mesh(nx, block, g, ns, **kw)
    d = BurgersDriver(m, cfl=cfl, recon=recon)
    om = OracleMesh(nx, block, g, boxes=kw.get("refine_boxes", ()),
                    spheres=kw.get("refine_spheres", ()))
    ob = OracleBurgers(om, ns, cfl, recon)
    return m, d, om, ob


def assert_close(got, exp, what):
    got = np.asarray(got).ravel()
    exp = np.asarray(exp).ravel()
    scale = np.max(np.abs(exp))
    err = np.max(np.abs(got - exp)) / scale if scale > 0 else np.max(np.abs(got - exp))
    print(f"{what}: max rel err {err:.3e}, bit-exact {np.array_equal(got, exp)}", flush=True)
    assert err <= RTOL, what


def assert_clock(dt, t, ob, exact):
    """dt and time against the oracle: bit for bit for the forms built without FMA contraction,
    within RTOL for the block sweep (built with contraction, like a CUDA build of the
    reference)."""
    from oracle_lib import lib
    t_exp = lib().orc_burgers_time(ob.h)
    if exact:
        assert dt == ob.dt and t == t_exp
    else:
        assert abs(dt - ob.dt) <= RTOL * ob.dt and abs(t - t_exp) <= RTOL * t_exp, (dt, ob.dt)


@pytest.mark.parametrize("recon", [0, 1], ids=["weno5", "linear"])
def test_stage_kernels_match_oracle(recon):
    print(f"case stage kernels recon={recon}", flush=True)
    m, d, om, ob = pair(recon=recon)
    from oracle_lib import lib
    lib().orc_burgers_pgen_sinusoid(ob.h, 11)
    lib().orc_burgers_initialize(ob.h)
    d.initialize("sinusoid", seed=11)
    U = d.base.download("U")
    assert np.array_equal(U.ravel(), ob.get(0)), "initial state incl. ghosts"
    assert d.dt == ob.dt
    pa.burgers_package.CalculateFluxes(d.base, recon)
    lib().orc_burgers_fluxes(ob.h)
    F = d.base.download("U", which=2)
    for dim in range(3):
        assert_close(F[dim], ob.get(4 + dim), f"flux x{dim + 1}")
    pa.Update.FluxDivergence(d.base, d.dudt)
    lib().orc_burgers_fluxdiv(ob.h)
    assert_close(d.dudt.download("U"), ob.get(3), "dudt")
    pa.burgers_package.CalculateDerived(d.base)
    lib().orc_burgers_derived(ob.h)
    assert_close(d.base.download("derived"), ob.get(2), "derived")
    dt = pa.burgers_package.EstimateTimestepMesh(d.base, 0.4)
    assert dt == lib().orc_burgers_estimate_dt(ob.h)


@pytest.mark.parametrize("pgen", ["sinusoid", "vibe"])
def test_ten_cycles_match_oracle(pgen):
    # 64^3 mesh, 64 blocks of 16^3, nghost 4, 3 velocities + 4 scalars, WENO5, CFL 0.4
    print(f"case 10 cycles pgen={pgen}", flush=True)
    from oracle_lib import lib
    m, d, om, ob = pair(nx=(64, 64, 64))
    if pgen == "sinusoid":
        lib().orc_burgers_pgen_sinusoid(ob.h, 0)
    else:
        lib().orc_burgers_pgen_vibe(ob.h)
    lib().orc_burgers_initialize(ob.h)
    d.initialize(pgen, seed=0)
    for c in range(10):
        d.cycle()
    lib().orc_burgers_cycles(ob.h, 10)
    assert_close(d.base.download("U"), ob.get(0), "U after 10 cycles")
    assert abs(d.dt - ob.dt) <= RTOL * ob.dt
    assert abs(d.time - lib().orc_burgers_time(ob.h)) <= RTOL * d.time


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_device_driver_matches_oracle_every_cycle(use_graph, form):
    # fused stage launches, device-resident clock; with the graph, one captured cycle replayed,
    # with host downloads (eager work) between the replays, each replay checked; with overlap,
    # each stage's exchange runs on a second stream beside the ghost-free reconstruction
    print(f"case device driver graph={use_graph} form={form}", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    from oracle_lib import lib
    nx, block, g, ns = (64, 64, 64), (16, 16, 16), 4, 4
    m = make_burgers_mesh(nx, block, g, ns)
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, **FORMS[form])
    d.initialize("sinusoid", seed=3)
    ob = OracleBurgers(OracleMesh(nx, block, g), ns, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 3)
    lib().orc_burgers_initialize(ob.h)
    dt, t = d.clock()
    assert dt == ob.dt and t == 0.0
    for c in range(4):
        d.cycles(1)
        lib().orc_burgers_cycles(ob.h, 1)
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert_close(d.base.download("derived"), ob.get(2), f"derived after cycle {c + 1}")
        assert_clock(*d.clock(), ob, exact=form != "sweep")
    d.cycles(6)
    lib().orc_burgers_cycles(ob.h, 6)
    assert_close(d.base.download("U"), ob.get(0), "U after 10 cycles")


def test_cpp_task_driver_matches_oracle(tmp_path):
    # the reference's task order written in C++ against include/parthenon_amd.hpp
    import os
    import subprocess
    from oracle_lib import lib
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "burgers_task_driver")
    out = str(tmp_path / "u.bin")
    print("case cpp task driver 5 cycles", flush=True)
    r = subprocess.run([exe, "5", out], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    U, dt, t = raw[:-2], raw[-2], raw[-1]
    ob = OracleBurgers(OracleMesh((64, 64, 64), (16, 16, 16), 4), 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 0)
    lib().orc_burgers_initialize(ob.h)
    lib().orc_burgers_cycles(ob.h, 5)
    assert_close(U, ob.get(0), "C++ driver U after 5 cycles")
    assert dt == ob.dt and t == lib().orc_burgers_time(ob.h)


def test_device_driver_rejects_small_nghost():
    from parthenon_amd.burgers import DeviceBurgersDriver
    m = make_burgers_mesh((32, 32, 32), (16, 16, 16), 2, 1)
    d = DeviceBurgersDriver(m)
    with pytest.raises(RuntimeError, match="ghost"):
        d.initialize("sinusoid", seed=0)


def test_conservation_to_roundoff():
    print("case conservation", flush=True)
    m, d, om, ob = pair(nx=(32, 32, 32), block=(8, 8, 8))
    d.initialize("sinusoid", seed=5)
    g = m.ng[0]

    def totals():
        U = d.base.download("U")[:, :, g:-g, g:-g, g:-g]
        return U.sum(axis=(0, 2, 3, 4))

    t0 = totals()
    for _ in range(3):
        d.cycle()
    t1 = totals()
    n = 32 ** 3
    # periodic flux form: sum of U over the mesh changes only by round-off
    assert np.all(np.abs(t1 - t0) <= 1e-12 * n * np.maximum(1.0, np.abs(t0) / n)), (t0, t1)


# Static refinement: a box refined once on a 32^3 mesh of 8^3 blocks (120 blocks), and the
# static-refined benchmark mesh: 64^3 in 16^3 blocks, two levels on the blocks meeting the
# sphere of radius 0.25 at the centre (960 blocks), nghost 4.
REFINED = {
    "box120": dict(nx=(32, 32, 32), block=(8, 8, 8),
                   refine_boxes=[((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 1)]),
    "sphere960": dict(nx=(64, 64, 64), block=(16, 16, 16),
                      refine_spheres=[((0.0, 0.0, 0.0), 0.25, 2)]),
}


@pytest.mark.parametrize("name", ["box120", "sphere960"])
def test_multilevel_task_driver_matches_oracle(name):
    # task-by-task stage: fluxes, flux correction, divergence, update, exchange with
    # restriction and prolongation
    spec = REFINED[name]
    print(f"case multilevel task driver {name}", flush=True)
    from oracle_lib import lib
    m, d, om, ob = pair(nx=spec["nx"], block=spec["block"],
                        refine_boxes=spec.get("refine_boxes", ()),
                        refine_spheres=spec.get("refine_spheres", ()))
    assert m.multilevel and m.nblocks == om.nblocks
    lib().orc_burgers_pgen_sinusoid(ob.h, 2)
    lib().orc_burgers_initialize(ob.h)
    d.initialize("sinusoid", seed=2)
    assert np.array_equal(d.base.download("U").ravel(), ob.get(0)), "initial state"
    ncyc = 3 if name == "box120" else 2
    for c in range(ncyc):
        d.cycle()
        lib().orc_burgers_cycles(ob.h, 1)
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert d.dt == ob.dt


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("name", ["box120", "sphere960"])
def test_multilevel_device_driver_matches_oracle(name, use_graph, form):
    # fused stage with the flux-correction launches between reconstruction and update
    spec = REFINED[name]
    print(f"case multilevel device driver {name} graph={use_graph} form={form}", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    from oracle_lib import lib
    m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4,
                          refine_boxes=spec.get("refine_boxes", ()),
                          refine_spheres=spec.get("refine_spheres", ()))
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, **FORMS[form])
    d.initialize("sinusoid", seed=2)
    om = OracleMesh(spec["nx"], spec["block"], 4, boxes=spec.get("refine_boxes", ()),
                    spheres=spec.get("refine_spheres", ()))
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 2)
    lib().orc_burgers_initialize(ob.h)
    dt, _ = d.clock()
    assert dt == ob.dt
    ncyc = 3 if name == "box120" else 2
    for c in range(ncyc):
        d.cycles(1)
        lib().orc_burgers_cycles(ob.h, 1)
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert_clock(*d.clock(), ob, exact=form != "sweep")


def test_cpp_task_driver_multilevel_matches_oracle(tmp_path):
    import os
    import subprocess
    from oracle_lib import lib
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "burgers_task_driver")
    out = str(tmp_path / "u.bin")
    print("case cpp task driver sphere960 2 cycles", flush=True)
    r = subprocess.run([exe, "2", out, "0.25"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    U, dt, t = raw[:-2], raw[-2], raw[-1]
    om = OracleMesh((64, 64, 64), (16, 16, 16), 4, spheres=[((0.0, 0.0, 0.0), 0.25, 2)])
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 0)
    lib().orc_burgers_initialize(ob.h)
    lib().orc_burgers_cycles(ob.h, 2)
    assert_close(U, ob.get(0), "C++ driver U after 2 cycles (sphere960)")
    assert dt == ob.dt and t == lib().orc_burgers_time(ob.h)


def test_multilevel_conservation_to_roundoff():
    # with flux correction the two-level scheme conserves sum(U * cell volume) over the leaf
    # blocks to round-off
    print("case multilevel conservation", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    spec = REFINED["box120"]
    m = make_burgers_mesh(spec["nx"], spec["block"], 4, 1, refine_boxes=spec["refine_boxes"])
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=True)
    d.initialize("sinusoid", seed=7)
    g = m.ng[0]
    vol = np.array([np.prod(np.subtract(m.block(b)["xmax"], m.block(b)["xmin"])) / 8 ** 3
                    for b in range(m.nblocks)])

    def totals():
        U = d.base.download("U")[:, :, g:-g, g:-g, g:-g]
        return (U.sum(axis=(2, 3, 4)) * vol[:, None]).sum(axis=0)

    t0 = totals()
    d.cycles(4)
    t1 = totals()
    assert np.all(np.isfinite(t1))
    assert np.max(np.abs(t1 - t0)) < 1e-14, (t0, t1)


def test_history_octant_masses_match_oracle(tmp_path):
    # MassHistory on the refined mesh after a few cycles, written in the .hst format and
    # compared with the oracle's octant sums (Kokkos::Sum order is not fixed: 1e-12 relative)
    print("case history octant masses", flush=True)
    import os
    from oracle_lib import lib
    from parthenon_amd.burgers import DeviceBurgersDriver
    from parthenon_amd.history import HistoryOutput, rel_diff
    import ctypes as C
    spec = REFINED["box120"]
    m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4, refine_boxes=spec["refine_boxes"])
    d = DeviceBurgersDriver(m, cfl=0.4)
    d.initialize("sinusoid", seed=6)
    ob = OracleBurgers(OracleMesh(spec["nx"], spec["block"], 4, boxes=spec["refine_boxes"]),
                       4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 6)
    lib().orc_burgers_initialize(ob.h)
    hst = HistoryOutput(basename=str(tmp_path / "burgers"))
    mid, lo, hi = 0.0, -0.5, 0.5
    rows_o = []
    for c in range(3):
        d.cycles(1)
        lib().orc_burgers_cycles(ob.h, 1)
        dt, t = d.clock()
        got = hst.write(d.base, t, dt, c + 1, m.nblocks)
        exp = []
        for s1 in (0, 1):
            for s2 in (0, 1):
                for s3 in (0, 1):
                    xmn = [mid if s else lo for s in (s1, s2, s3)]
                    xmx = [hi if s else mid for s in (s1, s2, s3)]
                    exp.append(lib().orc_burgers_mass_history(
                        ob.h, (C.c_double * 3)(*xmn), (C.c_double * 3)(*xmx), 1.0))
        rows_o.append([t, dt, c + 1, m.nblocks] + exp)
        assert np.max(rel_diff(got, exp)) < 1e-12, (got, exp)
    table = np.loadtxt(hst.fname)
    assert table.shape == (3, 12)
    # the written table agrees with the oracle's rows at the printed precision
    assert np.max(rel_diff(table, np.array(rows_o))) < 1e-4
    with open(hst.fname) as f:
        assert f.readline().startswith("#  History data")


def test_tuning_choices_give_identical_bits():
    # the register budget of the block sweep and the plane split change the work split only:
    # the sweep's budgets give the same bits, the plane-tiled forms give the same bits, and the
    # two families (the sweep built with FMA contraction) agree within RTOL
    print("case tuning independence", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    spec = REFINED["box120"]
    outs = []
    for form, tune in (("sweep", [1, 1, 1, 4]), ("sweep", [1, 1, 1, 3]),
                       ("overlap", [2, 4, 2, 4]), ("serial", [4, 1, 1, 4])):
        m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4, refine_boxes=spec["refine_boxes"])
        d = DeviceBurgersDriver(m, cfl=0.4, use_graph=False, **FORMS[form])
        d.set_tuning(tune)
        d.initialize("sinusoid", seed=4)
        assert d.tuning() == tune
        d.cycles(2)
        outs.append(d.base.download("U"))
    # the automatic choice (whole cycles of each form timed on the saved state, which is then
    # restored) gives the bits of the form it picked
    m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4, refine_boxes=spec["refine_boxes"])
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=True)
    d.initialize("sinusoid", seed=4)
    print("automatic form:", d.form(), d.tuning(), flush=True)
    assert d.form() in FORMS
    d.cycles(2)
    auto = d.base.download("U")
    assert np.array_equal(outs[1], outs[0])
    assert np.array_equal(outs[3], outs[2])
    assert np.array_equal(auto, outs[0] if d.form() == "sweep" else outs[2])
    assert_close(outs[0], outs[2], "block sweep vs plane-tiled form")


@pytest.mark.parametrize("form", ["sweep", "serial"])
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_device_driver_time_limits_match_oracle(form, use_graph):
    # SetGlobalTimeStep's <parthenon/time> bounds and the end on tlim (driver.cpp:210-269) on
    # the device clock: a forced first dt, a dt ceiling, and tlim in the middle of the third
    # cycle; cycles queued after the end change nothing
    print(f"case device driver time limits form={form} graph={use_graph}", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    probe = OracleBurgers(OracleMesh((32, 32, 32), (16, 16, 16), 4), 4, 0.4, 0)
    probe.set_time_limits(dt_init=1e-4, dt_init_force=True, dt_ceil=6e-4)
    from oracle_lib import lib
    lib().orc_burgers_pgen_sinusoid(probe.h, 5)
    lib().orc_burgers_initialize(probe.h)
    probe.cycles(2)
    lim = dict(tlim=probe.time + 0.37 * probe.dt, dt_init=1e-4, dt_init_force=True,
               dt_ceil=6e-4)
    ob = OracleBurgers(OracleMesh((32, 32, 32), (16, 16, 16), 4), 4, 0.4, 0)
    ob.set_time_limits(**lim)
    lib().orc_burgers_pgen_sinusoid(ob.h, 5)
    lib().orc_burgers_initialize(ob.h)
    assert ob.cycles(10) == 3 and ob.time == lim["tlim"]
    m = make_burgers_mesh((32, 32, 32), (16, 16, 16), 4, 4)
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, **FORMS[form])
    d.set_time_limits(**lim)
    d.initialize("sinusoid", seed=5)
    assert d.clock()[0] == 1e-4
    d.cycles(6)
    assert_close(d.base.download("U"), ob.get(0), "U at tlim")
    assert_clock(*d.clock(), ob, exact=form != "sweep")


def test_sweep_form_sphere960_full_run_matches_oracle():
    # the benchmark's workload at its own length: the 960-block sphere mesh, 20 cycles of the
    # block-sweep form (built with FMA contraction), eager and graph-replayed side by side,
    # against the oracle at 1e-13 after cycles 1, 10 and 20
    print("case sphere960 sweep 20 cycles eager+graph", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    from oracle_lib import lib
    spec = REFINED["sphere960"]
    drivers = []
    for use_graph in (False, True):
        m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4,
                              refine_spheres=spec["refine_spheres"])
        d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, sweep=True)
        d.initialize("sinusoid", seed=2)
        drivers.append(d)
    om = OracleMesh(spec["nx"], spec["block"], 4, spheres=spec["refine_spheres"])
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 2)
    lib().orc_burgers_initialize(ob.h)
    done = 0
    for upto in (1, 10, 20):
        for d in drivers:
            d.cycles(upto - done)
        ob.cycles(upto - done)
        done = upto
        exp = ob.get(0)
        for d, name in zip(drivers, ("eager", "graph")):
            assert_close(d.base.download("U"), exp, f"sweep {name} U after {upto} cycles")
            assert_clock(*d.clock(), ob, exact=False)


@pytest.mark.parametrize("block", [(16, 16, 16), (8, 8, 8)], ids=["16cubed", "8cubed"])
def test_sweep_large_jump_matches_oracle(block):
    # WENO5Z's weights next to a jump of 1e30 in a passive scalar (flat on both sides): the
    # block sweep's regrouped reconstruction (Newton reciprocals of be0 be1 be2 and of each
    # side's S D) stays finite and agrees with the oracle's reference-order expression; the
    # velocities are O(1) and are checked on their own scale
    print(f"case large jump block={block}", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    from oracle_lib import lib
    m = make_burgers_mesh((32, 32, 32), block, 4, 4)
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=False, sweep=True)
    ob = OracleBurgers(OracleMesh((32, 32, 32), block, 4), 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 6)
    U = ob.get(0).reshape(m.field_shape("U"))
    g = m.ng[0]
    for b in range(m.nblocks):
        info = m.block(b)
        n = m.n[0]
        x = info["xmin"][0] + (np.arange(n) + 0.5) * (info["xmax"][0] - info["xmin"][0]) / n
        step = np.where(x > 0.0, 1.0e30, 0.0)
        U[b, 4, g:-g, g:-g, g:-g] = step[None, None, :]
    ob.set(0, U.ravel())
    lib().orc_burgers_initialize(ob.h)
    d.base.upload("U", U)
    d.initialize("none")
    d.cycles(2)
    ob.cycles(2)
    got = d.base.download("U")
    exp = ob.get(0).reshape(got.shape)
    assert np.isfinite(got).all()
    for c in range(got.shape[1]):
        assert_close(got[:, c], exp[:, c], f"component {c} after 2 cycles")


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_split_sweep_matches_unsplit(use_graph):
    # the split sweep (blocks without cross-level channels first, beside the deferred
    # cross-level exchange on a second stream) against the single-launch sweep: the same
    # operations on disjoint data; the two launches cut the planes into different workgroup
    # segments, and a segment that starts inside a block rebuilds its z state in code that the
    # FMA contraction of the sweep rounds differently, so the states agree to rounding (as
    # between the sweep's spread choices), ghosts and derived field included
    print(f"case split sweep sphere960 {'graph' if use_graph else 'eager'}", flush=True)
    from parthenon_amd.burgers import DeviceBurgersDriver
    spec = REFINED["sphere960"]
    drivers = []
    for split in (True, False):
        m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4,
                              refine_spheres=spec["refine_spheres"])
        d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, sweep=True, split=split)
        d.initialize("sinusoid", seed=5)
        drivers.append(d)
    done = 0
    for upto in (1, 5):
        for d in drivers:
            d.cycles(upto - done)
        done = upto
        a, b = drivers
        for name in ("U", "derived"):
            x, y = a.base.download(name), b.base.download(name)
            assert np.max(np.abs(x - y)) <= 1e-14 * np.max(np.abs(y)), (name, upto)
        (dta, ta), (dtb, tb) = a.clock(), b.clock()
        assert abs(dta - dtb) <= 1e-14 * dtb and abs(ta - tb) <= 1e-14 * tb
