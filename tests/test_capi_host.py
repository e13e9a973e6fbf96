"""Host-only checks of the C-ABI library (no GPU): it loads, exports every entry point the
header declares, and its host topology (blocks, neighbour links, region index boxes) equals the
oracle's restatement of the reference on uniform and refined meshes."""
import os
import re

import numpy as np
import pytest

import parthenon_amd as pa
from oracle_lib import OracleMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "parthenon_amd.h")).read()
    return sorted(set(re.findall(r"\b(pth_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = pa.lib()
    names = declared_entry_points()
    assert len(names) > 20
    for n in names:
        assert hasattr(L, n), n
    # and the Python mirror binds exactly those
    assert sorted(pa.exported_symbols()) == names


MESHES = [
    dict(nx=(32, 32, 32), block=(8, 8, 8), nghost=3),
    dict(nx=(16, 16, 16), block=(8, 8, 8), nghost=2),
    dict(nx=(64, 32, 16), block=(16, 16, 8), nghost=4),
    dict(nx=(32, 32, 32), block=(8, 8, 8), nghost=4,
         boxes=[((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 2)]),
    dict(nx=(64, 64, 64), block=(16, 16, 16), nghost=4, spheres=[((0.0, 0.0, 0.0), 0.25, 2)]),
    dict(nx=(32, 32, 32), block=(8, 8, 8), nghost=2,
         boxes=[((0.0, 0.2, -0.5, -0.3, 0.1, 0.4), 1), ((0.3, 0.5, 0.3, 0.5, -0.5, -0.4), 2)]),
]


def build_pair(spec):
    boxes = spec.get("boxes", ())
    spheres = spec.get("spheres", ())
    o = OracleMesh(spec["nx"], spec["block"], spec["nghost"], boxes=boxes, spheres=spheres)
    p = pa.Mesh(spec["nx"], spec["block"], spec["nghost"], refine_boxes=boxes,
                refine_spheres=spheres)
    p.plan()
    return o, p


@pytest.mark.parametrize("spec", MESHES, ids=lambda s: f"{s['nx']}/{s['block']}")
def test_topology_and_region_boxes_match_oracle(spec):
    o, p = build_pair(spec)
    assert p.nblocks == o.nblocks
    assert p.multilevel == o.multilevel
    assert (p.n, p.ng, p.cn, p.cng) == (o.n, o.ng, o.cn, o.cng)
    for gid in range(o.nblocks):
        bo, bp = o.block(gid), p.block(gid)
        assert bo["level"] == bp["level"] and bo["l"] == bp["l"]
        assert bo["xmin"] == bp["xmin"] and bo["xmax"] == bp["xmax"]
        no = sorted((n["gid"], n["level"], n["origin"], n["off"]) for n in o.neighbors(gid))
        npr = sorted((n["gid"], n["level"], n["origin"], n["off"]) for n in p.neighbors(gid))
        assert no == npr
        ro, rp = o.region_boxes(gid), p.region_boxes(gid)
        # both list regions per neighbour in their own neighbour order: compare as multisets
        assert sorted(map(tuple, ro)) == sorted(map(tuple, rp))


def test_exchange_counts_match_oracle():
    o, p = build_pair(dict(nx=(128, 128, 128), block=(16, 16, 16), nghost=3))
    p.add_field("q", 1)
    used, _ = o.exchange_counts([1])
    nbuf, nel = p.exchange_counts()
    assert nel == used == 512 * ((16 + 6) ** 3 - 16 ** 3)
    assert nbuf == 512 * 26


def test_cpp_interface_header_compiles_and_driver_links():
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "burgers_task_driver.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), src],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "burgers_task_driver"))


def test_refinement_rejects_odd_nghost():
    p = pa.Mesh((32, 32, 32), (8, 8, 8), 3, refine_boxes=[((-0.1, 0.1) * 3, 1)])
    with pytest.raises(RuntimeError, match="even nghost"):
        p.plan()


def test_bad_block_size_rejected():
    with pytest.raises(RuntimeError, match="divide"):
        pa.Mesh((30, 32, 32), (8, 8, 8), 2)
