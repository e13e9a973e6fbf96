"""GPU parity of adaptive mesh refinement through the C ABI against the oracle: the derivative
criteria of Refinement::Tag (tags bit-exact, probed with tolerances set to the oracle's own
maxima), the remesh (block list, restriction / copy / prolongation of every container and the
exchange of "base", bit-exact), and adaptive Burgers runs of both drivers (tag and remesh after
every cycle; U within 1e-13 relative, block lists exact, dt and time exact except for the block
sweep, which is built with FMA contraction: within 1e-13)."""
import numpy as np
import pytest

import parthenon_amd as pa
from parthenon_amd.burgers import BurgersDriver, DeviceBurgersDriver, make_burgers_mesh
from oracle_lib import OracleAmr, OracleBurgers, OracleMesh, lib

pytestmark = pytest.mark.gpu

RTOL = 1e-13
FORMS = {"sweep": dict(sweep=True), "overlap": dict(sweep=False, overlap=True),
         "serial": dict(sweep=False, overlap=False)}
# 32^3 cells in 8^3 blocks: 4^3 base blocks, root level 2; numlevel 3 lets the base blocks
# refine once (the criteria compare logical levels)
NX, BLOCK, G = (32, 32, 32), (8, 8, 8), 4
BOX = [((-0.5, -0.25, -0.5, -0.25, -0.5, -0.25), 1)]


def assert_close(got, exp, what):
    got, exp = np.asarray(got).ravel(), np.asarray(exp).ravel()
    assert got.shape == exp.shape, what
    scale = np.max(np.abs(exp))
    err = np.max(np.abs(got - exp)) / scale
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


def gpu_mesh(criteria=(), numlevel=3, derefine_count=1, boxes=(), ncomp=5, extra=False):
    m = pa.Mesh(NX, BLOCK, G, refine_boxes=boxes)
    m.add_field("q", ncomp, pa.Metadata.FillGhost | pa.Metadata.Independent)
    if extra:
        m.add_field("w", 2, pa.Metadata.OneCopy)
    m.set_adaptive(numlevel, derefine_count)
    for c in criteria:
        m.add_refinement_criterion("q", **c)
    m.finalize()
    return m


def seeded(shape, seed):
    rng = np.random.default_rng(seed)
    return 1.0 + rng.standard_normal(shape) * 0.3


def test_tags_match_oracle_bitwise():
    # Tolerances at the oracle's exact maxima: a block whose GPU maximum differed by one ulp
    # would flip between refine, keep and derefine.
    print("case amr tags", flush=True)
    om = OracleMesh(NX, BLOCK, G, boxes=BOX)
    oa = OracleAmr(om, 3, 1)
    q = seeded(om.field_shape(5), 3)
    om.exchange_bc([q], coarse=[np.zeros(om.coarse_shape(5))])
    for method, comp in ((1, 3), (2, 0)):
        mx = oa.derivative_max(q, comp=comp, method=method)
        for tol in (np.median(mx), mx[7], mx[om.nblocks - 1]):
            crit = [dict(comp=comp, method=method, refine_tol=float(tol), derefine_tol=float(tol))]
            m = gpu_mesh(crit, boxes=BOX)
            md = m.mesh_data("base")
            md.upload("q", q)
            pa.Refinement.Tag(md)
            oa2 = OracleAmr(om, 3, 1, criteria=[dict(method=method, **{k: v for k, v in
                                                     crit[0].items() if k != "method"})])
            oa2.tag(q)
            f, c = m.refinement_flags()
            fo, co = oa2.flags()
            print(f"method {method} tol {tol:.6g}: flags {np.bincount(f + 1, minlength=3)}",
                  flush=True)
            assert np.array_equal(f, fo) and np.array_equal(c, co)
            assert (f == 0).any()   # the block at the tolerance itself keeps its level


def test_remesh_moves_every_container_like_oracle():
    # Refine two base blocks and derefine the refined corner's sibling group in one update:
    # "base" is moved and exchanged, "1" and the OneCopy field are moved only.
    print("case amr remesh", flush=True)
    m = gpu_mesh(boxes=BOX, extra=True)
    om = OracleMesh(NX, BLOCK, G, boxes=BOX)
    oa = OracleAmr(om, 3, 1)
    assert m.locs() == om.locs()
    base, one = m.mesh_data("base"), m.mesh_data("1")
    qb, q1, w = seeded(m.field_shape("q"), 5), seeded(m.field_shape("q"), 6), seeded(
        m.field_shape("w"), 7)
    base.upload("q", qb)
    one.upload("q", q1)
    base.upload("w", w)
    locs = m.locs()
    refine = [g for g, (lv, l) in enumerate(locs) if lv == 0 and l in ((3, 3, 3), (2, 3, 3))]
    group = [g for g, (lv, l) in enumerate(locs) if lv == 1 and max(l) <= 1]
    assert len(refine) == 2 and len(group) == 8
    for g in range(m.nblocks):
        tag = 1 if g in refine else (-1 if g in group else 0)
        m.set_refinement(g, tag)
        oa.set_refinement(g, tag)
    f, c = m.refinement_flags()
    assert np.array_equal(f, oa.flags()[0]) and np.array_equal(c, oa.flags()[1])
    assert m.adapt()
    nb, nnew, ndel = oa.update()
    eb, e1, ew = oa.transfer(qb), oa.transfer(q1), oa.transfer(w)
    oa.commit()
    assert m.nblocks == om.nblocks == nb and m.locs() == om.locs()
    om.exchange_bc([eb], coarse=[np.zeros(om.coarse_shape(5))])
    assert np.array_equal(base.download("q"), eb), "base after remesh and exchange"
    assert np.array_equal(one.download("q"), e1), "stage container moved"
    assert np.array_equal(m.mesh_data("1").download("w"), ew), "OneCopy field moved"
    f, c = m.refinement_flags()
    assert np.array_equal(f, oa.flags()[0]) and np.array_equal(c, oa.flags()[1])
    # nothing flagged: the next adapt leaves the mesh alone
    for g in range(m.nblocks):
        m.set_refinement(g, 0)
    assert not m.adapt()


def test_remesh_round_trip_restores_data():
    print("case amr round trip", flush=True)
    m = gpu_mesh()
    md = m.mesh_data("base")
    q = seeded(m.field_shape("q"), 9)
    md.upload("q", q)
    pa.exchange_fused(md)
    q = md.download("q")
    for g in range(m.nblocks):
        m.set_refinement(g, 1 if g in (5, 6, 40) else 0)
    assert m.adapt() and m.nblocks > 64
    for g in range(m.nblocks):
        m.set_refinement(g, -1 if m.block(g)["level"] == 1 else 0)
    assert m.adapt() and m.nblocks == 64
    # interior of a refined-then-derefined block is the restriction of its prolongation:
    # compare against the oracle doing the same two steps
    om = OracleMesh(NX, BLOCK, G)
    oa = OracleAmr(om, 3, 1)
    x = q.copy()
    for tags in ("refine", "derefine"):
        for g in range(om.nblocks):
            lv = om.block(g)["level"]
            oa.set_refinement(g, (1 if g in (5, 6, 40) else 0) if tags == "refine"
                              else (-1 if lv == 1 else 0))
        oa.update()
        x = oa.transfer(x)
        oa.commit()
        om.exchange_bc([x], coarse=[np.zeros(om.coarse_shape(5))])
    assert np.array_equal(md.download("q"), x)


def test_adapt_errors():
    m = pa.Mesh(NX, BLOCK, G)
    m.add_field("q", 1)
    m.finalize()
    with pytest.raises(RuntimeError):
        m.adapt()
    m2 = gpu_mesh()
    with pytest.raises(RuntimeError):
        m2.set_refinement(m2.nblocks, 1)
    with pytest.raises(RuntimeError):
        m2.set_adaptive(2, 10)   # after finalize


VIBE_CRIT = [dict(field="U", comp=3, method="derivative_order_1", refine_tol=0.5,
                  derefine_tol=0.2)]


def oracle_adaptive(derefine_count):
    om = OracleMesh(NX, BLOCK, G)
    oa = OracleAmr(om, 3, derefine_count,
                   criteria=[dict(method=1, comp=3, refine_tol=0.5, derefine_tol=0.2)])
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 4)
    lib().orc_burgers_initialize(ob.h)
    return om, oa, ob


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_adaptive_device_driver_matches_oracle(form, use_graph):
    print(f"case amr device driver form={form} graph={use_graph}", flush=True)
    m = make_burgers_mesh(NX, BLOCK, G, 4, adaptive=dict(numlevel=3, derefine_count=2,
                                                         criteria=VIBE_CRIT))
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=use_graph, **FORMS[form])
    d.initialize("sinusoid", seed=4)
    om, oa, ob = oracle_adaptive(2)
    remeshed = 0
    for c in range(4):
        r = d.adaptive_cycles(1)
        ro = oa.burgers_cycles(ob, 1)
        remeshed += r
        assert r == ro and m.locs() == om.locs(), f"block list after cycle {c + 1}"
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert_clock(*d.clock(), ob, exact=form != "sweep")
        assert np.array_equal(m.refinement_flags()[0], oa.flags()[0])
    assert remeshed >= 1 and m.nblocks > 64


def test_adaptive_task_driver_matches_oracle():
    print("case amr task driver", flush=True)
    m = make_burgers_mesh(NX, BLOCK, G, 4, adaptive=dict(numlevel=3, derefine_count=2,
                                                         criteria=VIBE_CRIT))
    d = BurgersDriver(m, cfl=0.4)
    d.initialize("sinusoid", seed=4)
    om, oa, ob = oracle_adaptive(2)
    for c in range(3):
        r = d.adaptive_cycle()
        ro = oa.burgers_cycles(ob, 1)
        assert int(r) == ro and m.locs() == om.locs(), f"block list after cycle {c + 1}"
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert d.dt == ob.dt


def test_cpp_task_driver_adaptive_matches_oracle(tmp_path):
    # the reference's task order with Refinement::Tag and the remesh, in C++ against
    # include/parthenon_amd.hpp (64^3 in 16^3 blocks: root level 2, numlevel 3)
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "burgers_task_driver")
    out = str(tmp_path / "u.bin")
    print("case cpp task driver adaptive 3 cycles", flush=True)
    r = subprocess.run([exe, "3", out, "0", "periodic", "adaptive"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    U, dt, t = raw[:-2], raw[-2], raw[-1]
    om = OracleMesh((64, 64, 64), (16, 16, 16), 4)
    oa = OracleAmr(om, 3, 2, criteria=[dict(method=1, comp=3, refine_tol=0.5, derefine_tol=0.2)])
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 0)
    lib().orc_burgers_initialize(ob.h)
    assert oa.burgers_cycles(ob, 3) >= 1
    assert U.size == ob.get(0).size and om.nblocks > 64
    assert_close(U, ob.get(0), "C++ adaptive driver U after 3 cycles")
    assert dt == ob.dt and t == lib().orc_burgers_time(ob.h)


def test_remesh_2d_matches_oracle():
    # a 2-D mesh (nleaf 4): refine two blocks, then derefine one complete sibling group
    print("case amr 2-D remesh", flush=True)
    nx, block = (64, 64, 1), (16, 16, 1)
    m = pa.Mesh(nx, block, 2, periodic=(True, True, False))
    m.add_field("q", 3, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.set_adaptive(3, 1)
    m.finalize()
    om = OracleMesh(nx, block, 2, periodic=(1, 1, 0))
    oa = OracleAmr(om, 3, 1)
    md = m.mesh_data("base")
    q = seeded(m.field_shape("q"), 21)
    md.upload("q", q)
    pa.exchange_fused(md)
    x = md.download("q")
    for step in ("refine", "derefine"):
        locs = m.locs()
        assert locs == om.locs()
        for g, (lv, l) in enumerate(locs):
            if step == "refine":
                tag = 1 if l in ((1, 1, 0), (2, 2, 0)) else 0
            else:
                tag = -1 if lv == 1 and l[0] // 2 == 1 and l[1] // 2 == 1 else 0
            m.set_refinement(g, tag)
            oa.set_refinement(g, tag)
        assert m.adapt()
        nb, nnew, ndel = oa.update()
        x = oa.transfer(x)
        oa.commit()
        om.exchange_bc([x], coarse=[np.zeros(om.coarse_shape(3))])
        assert m.locs() == om.locs() and m.nblocks == nb
        assert np.array_equal(md.download("q"), x), f"2-D data after {step}"
    assert m.nblocks == 16 + 3


def test_adaptive_open_boundaries_match_oracle():
    # reflecting and outflow faces: the remesh's exchange applies the conditions (coarse
    # buffers before the prolongation), every cycle's state against the oracle
    print("case amr open boundaries", flush=True)
    bcs = ("reflect", "outflow", "periodic", "periodic", "outflow", "reflect")
    m = make_burgers_mesh(NX, BLOCK, G, 4, bcs=bcs,
                          adaptive=dict(numlevel=3, derefine_count=2, criteria=VIBE_CRIT))
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=True, sweep=True)
    d.initialize("sinusoid", seed=4)
    om = OracleMesh(NX, BLOCK, G, periodic=(0, 1, 0), bcs=bcs)
    oa = OracleAmr(om, 3, 2, criteria=[dict(method=1, comp=3, refine_tol=0.5, derefine_tol=0.2)])
    ob = OracleBurgers(om, 4, 0.4, 0)
    lib().orc_burgers_pgen_sinusoid(ob.h, 4)
    lib().orc_burgers_initialize(ob.h)
    remeshed = 0
    for c in range(3):
        r = d.adaptive_cycles(1)
        assert r == oa.burgers_cycles(ob, 1) and m.locs() == om.locs(), f"cycle {c + 1}"
        remeshed += r
        assert_close(d.base.download("U"), ob.get(0), f"U after cycle {c + 1}")
        assert_clock(*d.clock(), ob, exact=False)
    assert remeshed >= 1


def test_vibe_input_never_refines():
    # benchmarks/burgers/burgers.pin:21-55: 128^3 in 16^3 blocks (one tree, root level 3),
    # numlevel 2, derivative_order_1 on U[3] with 0.5 / 0.2. The criterion's max_level (2) is
    # below every logical level, so no block is tagged for refinement, and root-level blocks
    # never derefine: the mesh stays uniform, as in the oracle.
    print("case amr VIBE input", flush=True)
    m = make_burgers_mesh((128, 128, 128), (16, 16, 16), 4, 4,
                          adaptive=dict(numlevel=2, derefine_count=10, criteria=VIBE_CRIT))
    assert m.root_level == 3
    d = DeviceBurgersDriver(m, cfl=0.4, use_graph=True)
    d.initialize("vibe")
    assert d.adaptive_cycles(2) == 0 and m.nblocks == 512
    f, c = m.refinement_flags()
    assert not f.any() and not c.any()


def test_two_criteria_match_oracle():
    # first- and second-derivative criteria on different components, tolerances at the
    # oracle's exact maxima of each: the flags (max over criteria, then SetRefinement) agree
    print("case amr two criteria", flush=True)
    nx, block, box = (48, 48, 48), (16, 16, 16), [((-0.5, -0.2, -0.5, -0.2, -0.5, -0.2), 1)]
    om = OracleMesh(nx, block, 4, boxes=box)
    q = seeded(om.field_shape(2), 31)
    om.exchange_bc([q], coarse=[np.zeros(om.coarse_shape(2))])
    base = OracleAmr(om, 2, 1)
    m1 = base.derivative_max(q, comp=0, method=1)
    m2 = base.derivative_max(q, comp=1, method=2)
    for t1, t2 in ((np.median(m1), np.median(m2)), (m1.min(), m2.max()), (m1.max(), m2.min())):
        crit = [dict(method=1, comp=0, refine_tol=float(t1), derefine_tol=float(t1) * 0.9),
                dict(method=2, comp=1, refine_tol=float(t2), derefine_tol=float(t2) * 0.9)]
        m = pa.Mesh(nx, block, 4, refine_boxes=box)
        m.add_field("q", 2, pa.Metadata.FillGhost | pa.Metadata.Independent)
        m.set_adaptive(2, 1)
        for c in crit:
            m.add_refinement_criterion("q", **c)
        m.finalize()
        md = m.mesh_data("base")
        md.upload("q", q)
        pa.Refinement.Tag(md)
        oa = OracleAmr(om, 2, 1, criteria=crit)
        oa.tag(q)
        f, c = m.refinement_flags()
        print(f"tolerances {t1:.4g} {t2:.4g}: flags {np.bincount(f + 1, minlength=3)}", flush=True)
        assert np.array_equal(f, oa.flags()[0]) and np.array_equal(c, oa.flags()[1])


def test_remesh_1d_matches_oracle():
    # a 1-D mesh (nleaf 2): refine three blocks (two neighbours and one apart), then derefine
    # the pair whose parent has no finer neighbour
    print("case amr 1-D remesh", flush=True)
    nx, block = (64, 1, 1), (8, 1, 1)
    m = pa.Mesh(nx, block, 2, periodic=(True, False, False))
    m.add_field("q", 2, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.set_adaptive(3, 1)
    m.finalize()
    om = OracleMesh(nx, block, 2, periodic=(1, 0, 0))
    oa = OracleAmr(om, 3, 1)
    md = m.mesh_data("base")
    md.upload("q", seeded(m.field_shape("q"), 41))
    pa.exchange_fused(md)
    x = md.download("q")
    for step in ("refine", "derefine"):
        locs = m.locs()
        assert locs == om.locs()
        for g, (lv, l) in enumerate(locs):
            if step == "refine":
                tag = 1 if l[0] in (2, 3, 6) and lv == 0 else 0
            else:
                tag = -1 if lv == 1 and l[0] // 2 == 6 else 0
            m.set_refinement(g, tag)
            oa.set_refinement(g, tag)
        assert m.adapt()
        nb, nnew, ndel = oa.update()
        x = oa.transfer(x)
        oa.commit()
        om.exchange_bc([x], coarse=[np.zeros(om.coarse_shape(2))])
        assert m.locs() == om.locs() and m.nblocks == nb
        assert np.array_equal(md.download("q"), x), f"1-D data after {step}"
    assert m.nblocks == 8 + 2
