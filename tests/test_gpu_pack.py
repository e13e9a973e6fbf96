"""The pack interface (SparsePack / PackVariables / PackVariablesAndFluxes) from HIP code: the
reference's CalculateDerived and FluxDivergence written against parthenon::SparsePack
(tests/cpp/pack_kernels.hip, built with -ffp-contract=off) equal the oracle bit for bit, and
the library's own kernels; pack bounds, coordinates, caching, errors and the rebuild after a
remesh."""
import ctypes as C
import os

import numpy as np
import pytest

import parthenon_amd as pa
from parthenon_amd.burgers import BurgersDriver, make_burgers_mesh
from oracle_lib import OracleBurgers, OracleMesh, lib as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def kernels():
    path = os.path.join(HERE, "cpp", "libpack_kernels.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    L.pack_calculate_derived.argtypes = [C.c_void_p]
    L.pack_flux_divergence.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.pack_coordinates.argtypes = [C.c_void_p, C.c_char_p, C.c_int,
                                   C.POINTER(C.c_double), C.c_long]
    return L


def interior(a, g):
    return a[..., g:-g, g:-g, g:-g]


SPECS = {
    "uniform": dict(nx=(32, 32, 32), block=(16, 16, 16)),
    "sphere960": dict(nx=(64, 64, 64), block=(16, 16, 16),
                      refine_spheres=[((0.0, 0.0, 0.0), 0.25, 2)]),
}


def setup(name, seed):
    spec = SPECS[name]
    kw = {k: v for k, v in spec.items() if k not in ("nx", "block")}
    m = make_burgers_mesh(spec["nx"], spec["block"], 4, 4, **kw)
    d = BurgersDriver(m, cfl=0.4)
    d.initialize("sinusoid", seed=seed)
    om = OracleMesh(spec["nx"], spec["block"], 4, spheres=spec.get("refine_spheres", ()))
    ob = OracleBurgers(om, 4, 0.4, 0)
    orc().orc_burgers_pgen_sinusoid(ob.h, seed)
    orc().orc_burgers_initialize(ob.h)
    return m, d, ob


@pytest.mark.parametrize("name", list(SPECS))
def test_pack_kernels_match_oracle(name):
    print(f"case pack kernels {name}", flush=True)
    K = kernels()
    m, d, ob = setup(name, 7)
    g = m.ng[0]
    # CalculateDerived through the pack {"derived", "U"}: v(b, 4) is U component 3
    d.base.upload("derived", np.zeros(m.field_shape("derived")))
    assert K.pack_calculate_derived(d.base._h) == pa.TaskStatus.complete
    orc().orc_burgers_derived(ob.h)
    got = d.base.download("derived")
    exp = ob.get(2).reshape(got.shape)
    assert np.array_equal(interior(got, g), interior(exp, g))
    # FluxDivergence through PackVariablesAndFluxes("U") and PackVariables("U")
    pa.burgers_package.CalculateFluxes(d.base, 0)
    orc().orc_burgers_fluxes(ob.h)
    orc().orc_burgers_fluxdiv(ob.h)
    d.dudt.upload("U", np.zeros(m.field_shape("U")))
    assert K.pack_flux_divergence(d.base._h, d.dudt._h, 3) == pa.TaskStatus.complete
    got = d.dudt.download("U")
    assert np.array_equal(interior(got, g), interior(ob.get(3).reshape(got.shape), g))
    # and the library's own kernel gives the same bits
    pa.Update.FluxDivergence(d.base, d.dudt)
    assert np.array_equal(interior(d.dudt.download("U"), g), interior(got, g))


def test_pack_bounds_coordinates_and_cache():
    print("case pack bounds and coordinates", flush=True)
    K = kernels()
    m, d, ob = setup("sphere960", 3)
    h, v = d.base.pack(["derived", "U"])
    assert (v.nblocks, v.nvar, v.nfields) == (m.nblocks, 8, 2)
    assert d.base.pack_bounds(h, 0) == (0, 0) and d.base.pack_bounds(h, 1) == (1, 7)
    assert (v.is_, v.ie, v.ks, v.ke) == (4, 19, 4, 19) and tuple(v.ent) == (24, 24, 24)
    # cached: the same request gives the same pack; other options give another
    h2, v2 = d.base.pack(["derived", "U"])
    assert h2 == h and v2.data == v.data
    hf, vf = d.base.pack(["U"], pa.PDOpt.WithFluxes)
    assert hf != h and vf.flux and vf.nvar == 7
    # cell centres, fine and coarse bricks, against the block extents
    for coarse in (0, 1):
        hc, vc = d.base.pack(["U"], pa.PDOpt.Coarse if coarse else 0)
        ni, nj, nk = vc.ie - vc.is_ + 1, vc.je - vc.js + 1, vc.ke - vc.ks + 1
        n = m.nblocks * ni * nj * nk
        out = np.empty(3 * n)
        assert K.pack_coordinates(d.base._h, b"U", coarse,
                                  out.ctypes.data_as(C.POINTER(C.c_double)), 3 * n) == 0
        out = out.reshape(3, m.nblocks, nk, nj, ni)
        for b in (0, m.nblocks // 2, m.nblocks - 1):
            info = m.block(b)
            lo, hi = info["xmin"], info["xmax"]
            for dim, cnt in enumerate((ni, nj, nk)):
                dx = (hi[dim] - lo[dim]) / cnt
                exp = lo[dim] + (np.arange(cnt) + 0.5) * dx
                got = np.moveaxis(out[dim, b], 2 - dim, 0)[:, 0, 0]
                assert np.max(np.abs(got - exp)) < 1e-14, (coarse, b, dim)


def test_pack_errors():
    print("case pack errors", flush=True)
    m, d, ob = setup("uniform", 1)
    with pytest.raises(RuntimeError, match="no field"):
        d.base.pack(["nosuch"])
    with pytest.raises(RuntimeError, match="no fluxes"):
        d.base.pack(["derived"], pa.PDOpt.WithFluxes)
    with pytest.raises(RuntimeError, match="single-level"):
        d.base.pack(["U"], pa.PDOpt.Coarse)


def test_pack_rebuilt_after_remesh():
    # a remesh moves every brick: the next request rebuilds the pack over the new blocks
    print("case pack after remesh", flush=True)
    K = kernels()
    m = make_burgers_mesh((32, 32, 32), (16, 16, 16), 4, 4,
                          adaptive=dict(numlevel=2, derefine_count=10))
    base = m.mesh_data("base")
    pa.burgers_package.ProblemGeneratorSinusoid(base, 9)
    h0, v0 = base.pack(["derived", "U"])
    assert v0.nblocks == 8
    m.set_refinement(0, 1)
    assert m.adapt()
    h1, v1 = base.pack(["derived", "U"])
    assert v1.nblocks == 15 and v1.nvar == 8
    base.upload("derived", np.zeros(m.field_shape("derived")))
    assert K.pack_calculate_derived(base._h) == pa.TaskStatus.complete
    U = base.download("U")
    exp = 0.5 * U[:, 3] * (U[:, 0] * U[:, 0] + U[:, 1] * U[:, 1] + U[:, 2] * U[:, 2])
    g = m.ng[0]
    assert np.array_equal(interior(base.download("derived")[:, 0], g), interior(exp, g))
