"""GPU parity of the boundary exchange (SendBoundBufs / ReceiveBoundBufs / SetBounds /
ProlongateBounds) through the C ABI, against closed forms and the oracle. Ghost fill, pack,
unpack and restriction are bit-exact; prolongation is held to 1e-13 relative (it is in fact
bit-exact too, since both sides evaluate the same expression order without contraction)."""
import numpy as np
import pytest

import parthenon_amd as pa
from meshcheck import global_index_field, interior_mask, linear_field
from oracle_lib import OracleMesh
from parthenon_amd.inputs import interior_uniform

pytestmark = pytest.mark.gpu


def product_mesh(nx, block, g, nfields=1, ncomp=1, **kw):
    m = pa.Mesh(nx, block, g, **kw)
    for f in range(nfields):
        m.add_field(f"q{f}", ncomp, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.finalize()
    return m


@pytest.mark.parametrize("nx,block,g", [
    ((32, 32, 32), (8, 8, 8), 3),
    ((16, 16, 16), (8, 8, 8), 2),
    ((8, 8, 8), (8, 8, 8), 2),
    ((32, 16, 8), (8, 8, 8), 3),
])
def test_ghosts_equal_wrapped_interior(nx, block, g):
    print(f"case nx={nx} block={block} g={g}", flush=True)
    m = product_mesh(nx, block, g, ncomp=2)
    md = m.mesh_data("base")
    exp = global_index_field(m, 2, nx)
    md.upload("q0", np.where(interior_mask(m)[None, None], exp, np.nan))
    pa.exchange(md)
    assert np.array_equal(md.download("q0"), exp)


@pytest.mark.parametrize("mode", ["persistent", "per_buffer", "fused"])
def test_exchange_microbenchmark_shape_matches_oracle(mode):
    # 128^3 mesh in 512 blocks of 16^3, nghost 3, 8 one-component fields, U(-1,1) interiors
    print(f"case 128^3/16^3 g=3 x8 mode={mode}", flush=True)
    nx, block, g = (128, 128, 128), (16, 16, 16), 3
    m = product_mesh(nx, block, g, nfields=8)
    m.set_pack_mode(pa.PackMode.per_buffer if mode == "per_buffer" else pa.PackMode.persistent)
    md = m.mesh_data("base")
    o = OracleMesh(nx, block, g)
    fields = [interior_uniform(m.field_shape(f"q{f}"), m.ng, seed=1000 + f, fill=0.0)
              for f in range(8)]
    for f, a in enumerate(fields):
        md.upload(f"q{f}", a)
    if mode == "fused":
        pa.exchange_fused(md)
    else:
        pa.exchange(md)
    o.exchange(fields)
    exp = global_index_field(m, 1, nx)
    for f in range(8):
        got = md.download(f"q{f}")
        assert np.array_equal(got, fields[f]), f
    # second exchange on unchanged data is idempotent
    pa.exchange(md)
    assert np.array_equal(md.download("q3"), fields[3])
    del exp


def test_task_state_machine():
    m = product_mesh((16, 16, 16), (8, 8, 8), 2)
    md = m.mesh_data("base")
    assert pa.ReceiveBoundBufs(md) == pa.TaskStatus.incomplete
    with pytest.raises(RuntimeError, match="received"):
        pa.SetBounds(md)
    assert pa.SendBoundBufs(md) == pa.TaskStatus.complete
    assert pa.SendBoundBufs(md) == pa.TaskStatus.incomplete  # previous exchange unfinished
    assert pa.ReceiveBoundBufs(md) == pa.TaskStatus.complete
    assert pa.SetBounds(md) == pa.TaskStatus.complete
    assert pa.SendBoundBufs(md, pa.BoundaryType.nonlocal_) == pa.TaskStatus.complete
    assert pa.ReceiveBoundBufs(md, pa.BoundaryType.nonlocal_) == pa.TaskStatus.complete


# refinement kept away from the periodic boundary, where a linear field is not periodic
REFINED = [
    dict(nx=(32, 32, 32), block=(8, 8, 8), g=4, boxes=[((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 2)]),
    dict(nx=(64, 64, 64), block=(8, 8, 8), g=2,
         boxes=[((0.0, 0.2, -0.1, 0.1, 0.1, 0.3), 1), ((-0.2, -0.1, -0.2, -0.1, -0.2, -0.1), 2)]),
]


@pytest.mark.parametrize("spec", REFINED, ids=["box2", "two-boxes"])
def test_multilevel_linear_field_closed_form(spec):
    print(f"case refined {spec}", flush=True)
    m = product_mesh(spec["nx"], spec["block"], spec["g"], ncomp=2,
                     refine_boxes=spec["boxes"])
    assert m.multilevel
    md = m.mesh_data("base")
    coeffs = [(3.0, 2.0, 4.0, -8.0), (-1.0, 0.5, -0.25, 1.0)]
    exp = linear_field(m, 2, coeffs)
    md.upload("q0", np.where(interior_mask(m)[None, None], exp, np.nan))
    md.upload("q0", np.full(m.coarse_shape("q0"), np.nan), which=1)
    pa.exchange(md)
    assert np.array_equal(md.download("q0"), exp)


@pytest.mark.parametrize("spec", REFINED + [
    dict(nx=(64, 64, 64), block=(16, 16, 16), g=4, spheres=[((0.0, 0.0, 0.0), 0.25, 2)])],
    ids=["box2", "two-boxes", "sphere960"])
@pytest.mark.parametrize("fused", [False, True], ids=["tasks", "fused"])
def test_multilevel_random_data_matches_oracle(spec, fused):
    print(f"case refined random {spec} fused={fused}", flush=True)
    boxes, spheres = spec.get("boxes", ()), spec.get("spheres", ())
    m = product_mesh(spec["nx"], spec["block"], spec["g"], ncomp=3, refine_boxes=boxes,
                     refine_spheres=spheres)
    o = OracleMesh(spec["nx"], spec["block"], spec["g"], boxes=boxes, spheres=spheres)
    md = m.mesh_data("base")
    f = interior_uniform(m.field_shape("q0"), m.ng, seed=7, fill=0.0)
    c = np.zeros(m.coarse_shape("q0"))
    md.upload("q0", f)
    # coarse buffers start as NaN: a cell the prolongation reads that the restriction skipped
    # would turn the prolongated ghosts into NaN
    md.upload("q0", np.full(m.coarse_shape("q0"), np.nan), which=1)
    if fused:
        pa.exchange_fused(md)
    else:
        pa.exchange(md)
    o.exchange([f], coarse=[c], prolong=1)
    got, gotc = md.download("q0"), md.download("q0", which=1)
    # restriction into coarse buffers: bit-exact on every cell written. On these periodic
    # meshes the restriction after the set covers only the prolongation stencils' reach, a
    # subset of the reference's (and the oracle's) halo restriction.
    w = ~np.isnan(gotc)
    assert np.array_equal(gotc[w], c[w])
    print(f"coarse cells written: {w.mean():.3f} of the buffer", flush=True)
    mask = interior_mask(m)
    # copies into ghosts: bit-exact; prolongated ghosts within 1e-13 relative
    np.testing.assert_allclose(got, f, rtol=1e-13, atol=0)
    assert np.array_equal(got[:, :, mask], f[:, :, mask])
    bit_exact = np.array_equal(got, f)
    print("prolongation bit-exact:", bit_exact, flush=True)


@pytest.mark.parametrize("fused", [False, True], ids=["tasks", "fused"])
def test_registered_linear_prolongation_per_field(fused):
    # Metadata::RegisterRefinementOps per field: "lin" registers ProlongateSharedLinear, "mm"
    # keeps ProlongateSharedMinMod; one exchange serves both, each field against the oracle's
    # exchange with its own operator (the oracle's operators are pinned by the parabola
    # closed forms in test_oracle_exchange.py)
    print(f"case per-field prolongation fused={fused}", flush=True)
    spec = dict(nx=(64, 64, 64), block=(16, 16, 16), g=4)
    spheres = [((0.0, 0.0, 0.0), 0.25, 2)]
    m = pa.Mesh(spec["nx"], spec["block"], spec["g"], refine_spheres=spheres)
    for name in ("mm", "lin"):
        m.add_field(name, 2, pa.Metadata.FillGhost | pa.Metadata.Independent)
    m.register_refinement_ops("lin", "linear")
    m.finalize()
    md = m.mesh_data("base")
    results = {}
    for name, prolong, seed in (("mm", 1, 3), ("lin", 2, 4)):
        f = interior_uniform(m.field_shape(name), m.ng, seed=seed, fill=0.0)
        md.upload(name, f)
        o = OracleMesh(spec["nx"], spec["block"], spec["g"], spheres=spheres)
        c = np.zeros(m.coarse_shape(name))
        o.exchange([f], coarse=[c], prolong=prolong)
        results[name] = f
    (pa.exchange_fused if fused else pa.exchange)(md)
    for name in ("mm", "lin"):
        got = md.download(name)
        np.testing.assert_allclose(got, results[name], rtol=1e-13, atol=0)
        print(f"{name}: prolongation bit-exact {np.array_equal(got, results[name])}", flush=True)
    # the two operators really differ on this data
    f = interior_uniform(m.field_shape("mm"), m.ng, seed=4, fill=0.0)
    o = OracleMesh(spec["nx"], spec["block"], spec["g"], spheres=spheres)
    o.exchange([f], coarse=[np.zeros(m.coarse_shape("mm"))], prolong=1)
    assert not np.allclose(f, results["lin"], rtol=1e-13, atol=0)
