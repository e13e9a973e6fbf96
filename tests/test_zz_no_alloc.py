"""The allocation-free CFL estimate, pth_burgers_estimate_timestep_noalloc, through the C ABI:
its device-memory result equals the synchronous pth_burgers_estimate_timestep bit for bit
(the same min reduction and the same cfl product), on two calls in a row with the second on a
smaller mesh (the static scratch keeps the first call's partials beyond the second's), then
the first mesh again; repeated calls leave the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

import parthenon_amd as pa
from parthenon_amd.burgers import burgers_package, make_burgers_mesh

pytestmark = pytest.mark.gpu


def _noalloc(L):
    f = L.pth_burgers_estimate_timestep_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    return f


def _free_bytes(L):
    free, total = C.c_int64(), C.c_int64()
    pa._check(L.pth_device_memory(C.byref(free), C.byref(total)), "device_memory")
    return free.value


class _DeviceDoubles:
    """n doubles of device memory from the HIP runtime the library already uses (the test
    allocates them; the entry point under test does not)."""

    def __init__(self, n):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), 8 * n) == 0
        self.fill(float("nan"))

    def fill(self, v):
        h = np.full(self.n, v)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data, 8 * self.n, 1) == 0  # H2D

    def at(self, i):
        return C.c_void_p(self.ptr.value + 8 * i)

    def read(self):
        h = np.empty(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # D2H
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def _case(nx, seed):
    m = make_burgers_mesh(nx, (16, 16, 16), 4, 4)
    base = m.mesh_data("base")
    burgers_package.ProblemGeneratorSinusoid(base, seed)
    return m, base


def test_noalloc_dt_matches_synchronous_form():
    L = pa.lib()
    noalloc = _noalloc(L)
    cfl = 0.4
    big, big_md = _case((64, 64, 32), 5)     # 32 blocks: 512 reduction partials
    small, small_md = _case((32, 16, 16), 6)  # 2 blocks: 32 partials
    out = _DeviceDoubles(4)
    try:
        sync = []
        for step, (m, md, slot) in enumerate([(big, big_md, 0), (small, small_md, 1),
                                              (big, big_md, 2)]):
            print(f"case noalloc dt step {step} blocks {m.nblocks}", flush=True)
            dt = C.c_double()
            pa._check(L.pth_burgers_estimate_timestep(md._h, cfl, C.byref(dt)),
                      "estimate_timestep")
            pa._check(noalloc(md._h, cfl, out.at(slot)), "estimate_timestep_noalloc")
            pa._check(L.pth_mesh_synchronize(m._h), "synchronize")
            got = out.read()[slot]
            print(f"  synchronous {dt.value!r} noalloc {got!r}", flush=True)
            assert got == dt.value
            sync.append(dt.value)
        assert sync[0] == sync[2] and sync[1] != sync[0]

        # no device allocation: free memory is the same before and after repeated calls
        pa._check(L.pth_mesh_synchronize(big._h), "synchronize")
        before = _free_bytes(L)
        for _ in range(20):
            pa._check(noalloc(big_md._h, cfl, out.at(3)), "noalloc")
            pa._check(noalloc(small_md._h, cfl, out.at(3)), "noalloc")
        pa._check(L.pth_mesh_synchronize(big._h), "synchronize")
        assert _free_bytes(L) == before
        assert out.read()[3] == sync[1]
    finally:
        out.free()


def test_noalloc_rejects_null_output():
    L = pa.lib()
    m, md = _case((32, 16, 16), 7)
    assert _noalloc(L)(md._h, 0.4, None) == pa.TaskStatus.fail
    assert b"dt_device" in L.pth_last_error()
