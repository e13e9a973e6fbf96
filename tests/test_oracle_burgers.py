"""Pins the oracle's Burgers stencils against what the mathematics fixes, before it is trusted as
the checker of the HIP path: polynomial exactness of the WENO5Z and MC-linear reconstructions,
boundedness next to a discontinuity, mirror symmetries of the reconstruction and of the HLL
flux, exact steady states, the closed-form CFL time step and discrete conservation."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleBurgers, OracleMesh, dptr, lib


def weno(q):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 5)
    ql = np.empty(len(q))
    qr = np.empty(len(q))
    lib().orc_weno5z(len(q), dptr(q), dptr(ql), dptr(qr))
    return ql, qr


def linear(q):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    ql = np.empty(len(q))
    qr = np.empty(len(q))
    lib().orc_linear(len(q), dptr(q), dptr(ql), dptr(qr))
    return ql, qr


def hll(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 8)
    out = np.empty((len(a), 5))
    lib().orc_lr_to_flux(len(a), dptr(a), dptr(out))
    return out


rng = np.random.default_rng(3)


def test_weno_constant_data():
    c = rng.uniform(-5, 5, 200)
    ql, qr = weno(np.repeat(c, 5))
    np.testing.assert_allclose(ql, c, rtol=2e-15)
    np.testing.assert_allclose(qr, c, rtol=2e-15)


def test_weno_linear_data():
    a, b = rng.uniform(-3, 3, 200), rng.uniform(-2, 2, 200)
    q = a[:, None] + b[:, None] * np.arange(5)[None]
    ql, qr = weno(q)
    np.testing.assert_allclose(ql, q[:, 2] + b / 2, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(qr, q[:, 2] - b / 2, rtol=1e-13, atol=1e-14)


def test_weno_quadratic_cell_averages_give_exact_interface_values():
    # cell averages of x^2 over [i-1/2, i+1/2] are i^2 + 1/12; with equal smoothness
    # indicators the weights are ideal and the 5-point reconstruction is exact for degree 2
    for s in (-2.0, 0.0, 3.0):
        x = np.arange(5) + s
        ql, qr = weno(x ** 2 + 1.0 / 12.0)
        assert abs(ql[0] - (x[2] + 0.5) ** 2) < 1e-12
        assert abs(qr[0] - (x[2] - 0.5) ** 2) < 1e-12


def test_weno_bounded_next_to_a_jump():
    # a flat stencil next to a jump: smoothness indicators of 0 and O(100) must not overflow
    cases = np.array([[1, 1, 1, 1, 11], [11, 1, 1, 1, 1], [1, 1, 1, 11, 11], [0, 0, 0, 0, 1e-3],
                      [5, 5, 5, 5, 5 + 1e-12]], dtype=np.float64)
    ql, qr = weno(cases)
    assert np.all(np.isfinite(ql)) and np.all(np.isfinite(qr))
    lo, hi = cases.min(axis=1), cases.max(axis=1)
    assert np.all(ql >= lo - 1e-12) and np.all(ql <= hi + 1e-12)
    assert np.all(qr >= lo - 1e-12) and np.all(qr <= hi + 1e-12)


def test_weno_mirror_symmetry():
    q = rng.normal(size=(500, 5))
    ql, qr = weno(q)
    ql2, qr2 = weno(q[:, ::-1].copy())
    np.testing.assert_allclose(ql2, qr, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(qr2, ql, rtol=1e-14, atol=1e-15)


def test_linear_reconstruction():
    a, b = rng.uniform(-3, 3, 100), rng.uniform(-2, 2, 100)
    q = a[:, None] + b[:, None] * np.arange(3)[None]
    ql, qr = linear(q)
    np.testing.assert_allclose(ql, q[:, 1] + b / 2, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(qr, q[:, 1] - b / 2, rtol=1e-14, atol=1e-15)
    # a local extremum keeps the cell value on both faces
    ql, qr = linear(np.array([[0.0, 1.0, 0.5], [2.0, -1.0, 3.0]]))
    assert np.array_equal(ql, [1.0, -1.0]) and np.array_equal(qr, [1.0, -1.0])


def test_hll_mirror_symmetry_and_rest_state():
    a = rng.normal(size=(400, 8))
    f = hll(a)
    # reflect the face: swap left/right and negate every velocity; momentum fluxes are even
    m = np.empty_like(a)
    m[:, 0], m[:, 1] = -a[:, 1], -a[:, 0]
    m[:, 2], m[:, 3] = -a[:, 3], -a[:, 2]
    m[:, 4], m[:, 5] = -a[:, 5], -a[:, 4]
    m[:, 6], m[:, 7] = -a[:, 7], -a[:, 6]
    g = hll(m)
    assert np.array_equal(g[:, 0], -f[:, 1]) and np.array_equal(g[:, 1], -f[:, 0])
    np.testing.assert_array_equal(g[:, 2:], f[:, 2:])
    z = hll(np.zeros((1, 8)))
    assert np.array_equal(z, np.zeros((1, 5)))
    # signal speeds bracket zero and the normal velocities
    assert np.all(f[:, 0] <= 0) and np.all(f[:, 1] >= 0)
    assert np.all(f[:, 0] <= np.minimum(a[:, 6], a[:, 7]))
    assert np.all(f[:, 1] >= np.maximum(a[:, 6], a[:, 7]))


def burgers(nx=(32, 32, 32), block=(16, 16, 16), g=4, ns=2, cfl=0.4):
    m = OracleMesh(nx, block, g)
    return m, OracleBurgers(m, ns, cfl, 0)


def test_uniform_state_is_steady_and_dt_is_closed_form():
    m, b = burgers()
    U = b.get(0).reshape(m.field_shape(b.V))
    vel = (0.75, -0.5, 0.25)
    for c in range(b.V):
        U[:, c] = vel[c] if c < 3 else 2.0
    b.set(0, U)
    lib().orc_burgers_initialize(b.h)
    dx = 1.0 / 32
    assert b.dt == 0.4 * (1.0 / (0.75 / dx + 0.5 / dx + 0.25 / dx))
    lib().orc_burgers_cycles(b.h, 2)
    assert np.array_equal(b.get(0).reshape(U.shape), U)


def test_sinusoid_run_conserves_totals():
    m, b = burgers(nx=(32, 32, 32), block=(8, 8, 8))
    lib().orc_burgers_pgen_sinusoid(b.h, 9)
    lib().orc_burgers_initialize(b.h)
    g = m.ng[0]
    tot = lambda: b.get(0).reshape(m.field_shape(b.V))[:, :, g:-g, g:-g, g:-g].sum(axis=(0, 2, 3, 4))
    t0 = tot()
    lib().orc_burgers_cycles(b.h, 5)
    t1 = tot()
    assert np.all(np.isfinite(t1))
    assert np.max(np.abs(t1 - t0)) < 1e-10 * 32 ** 3


def test_vibe_problem_stays_finite():
    m, b = burgers(nx=(64, 64, 64), ns=4)
    lib().orc_burgers_pgen_vibe(b.h)
    lib().orc_burgers_initialize(b.h)
    lib().orc_burgers_cycles(b.h, 3)
    assert np.all(np.isfinite(b.get(0)))


def refined_burgers(flux_correction, cycles):
    """Two-level mesh: 32^3 cells in 8^3 blocks with the central 2^3 base blocks refined once."""
    m = OracleMesh((32, 32, 32), (8, 8, 8), 4, boxes=[((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 1)])
    b = OracleBurgers(m, 1, 0.4, 0)
    lib().orc_burgers_set_flux_correction(b.h, int(flux_correction))
    lib().orc_burgers_pgen_sinusoid(b.h, 21)
    lib().orc_burgers_initialize(b.h)
    t0 = b.totals()
    lib().orc_burgers_cycles(b.h, cycles)
    return m, b, t0, b.totals()


def test_flux_correction_makes_two_level_run_conservative():
    # Sum of U * cell volume over the leaf blocks changes only through unmatched fluxes across
    # coarse/fine faces: with flux correction it is conserved to round-off, without it is not.
    m, b, t0, t1 = refined_burgers(True, 3)
    assert m.multilevel and m.nblocks == 120
    assert np.max(np.abs(t1 - t0)) < 1e-14
    _, _, u0, u1 = refined_burgers(False, 3)
    assert np.array_equal(t0, u0)
    assert np.max(np.abs(u1 - u0)) > 1e-8


def coarse_fine_face_errors(m, b):
    """For every coarse face cell on a coarse/fine face, |coarse flux - mean of the 4 fine fluxes
    on the same face area|, matched by block geometry (not by the oracle's index boxes)."""
    g, n = m.ng[0], m.n[0]
    E = m.ent
    F = [b.get(4 + d).reshape(m.nblocks, b.V, E[2], E[1], E[0]) for d in range(3)]
    blocks = [m.block(i) for i in range(m.nblocks)]
    errs = []
    for gid, bk in enumerate(blocks):
        for nb in m.neighbors(gid):
            off = nb["off"]
            if nb["level"] != bk["level"] + 1 or sum(o != 0 for o in off) != 1:
                continue
            d = [t for t in range(3) if off[t]][0]
            fb = blocks[nb["gid"]]
            tr = [t for t in range(3) if t != d]
            dxc = [(bk["xmax"][t] - bk["xmin"][t]) / n for t in range(3)]
            dxf = [(fb["xmax"][t] - fb["xmin"][t]) / n for t in range(3)]
            for a in range(n):
                for c in range(n):
                    cen = {tr[0]: bk["xmin"][tr[0]] + (a + 0.5) * dxc[tr[0]],
                           tr[1]: bk["xmin"][tr[1]] + (c + 0.5) * dxc[tr[1]]}
                    if not all(fb["xmin"][t] < cen[t] < fb["xmax"][t] for t in tr):
                        continue
                    ci = {d: g if off[d] < 0 else g + n, tr[0]: g + a, tr[1]: g + c}
                    lo = {t: g + int(round((cen[t] - fb["xmin"][t]) / dxf[t])) - 1 for t in tr}
                    fine = []
                    for p in range(2):
                        for q in range(2):
                            fi = {d: g + n if off[d] < 0 else g, tr[0]: lo[tr[0]] + p,
                                  tr[1]: lo[tr[1]] + q}
                            fine.append(F[d][nb["gid"], :, fi[2], fi[1], fi[0]])
                    coarse = F[d][gid, :, ci[2], ci[1], ci[0]]
                    errs.append(np.max(np.abs(coarse - np.mean(fine, axis=0))))
    return np.array(errs)


def test_flux_correction_overwrites_each_coarse_face_with_the_fine_face_average():
    m, b, _, _ = refined_burgers(True, 1)
    e = coarse_fine_face_errors(m, b)
    # 8 refined base blocks expose 24 coarse/fine block faces of 8x8 coarse face cells
    assert len(e) == 24 * 64
    assert e.max() < 1e-14
    m, b, _, _ = refined_burgers(False, 1)
    assert coarse_fine_face_errors(m, b).max() > 1e-6


def qflux(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)
    out = np.empty(len(a))
    lib().orc_scalar_flux(len(a), dptr(a), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def textbook_hll(fl, fr, ul, ur, sl, sr):
    """Harten-Lax-van Leer flux for sl < 0 < sr (Toro, Riemann Solvers and Numerical Methods
    for Fluid Dynamics, 3rd ed., eq. 10.21): (sr F_L - sl F_R + sl sr (U_R - U_L)) / (sr - sl)."""
    return (sr * fl - sl * fr + sl * sr * (ur - ul)) / (sr - sl)


def straddling_states(n, seed):
    r = np.random.default_rng(seed)
    a = r.uniform(-2, 2, size=(n, 8))
    # normal velocities of opposite signs, in either order, so that sl < 0 < sr
    up = r.uniform(0.05, 2, size=(n, 2)) * np.where(r.random((n, 1)) < 0.5, [1, -1], [-1, 1])
    a[:, 6:8] = up
    return a


def test_hll_is_the_textbook_flux_when_the_speeds_straddle_zero():
    # burgers_package.hpp:33-43 and burgers_package.cpp:335-340: the Davis speeds sl = min(upl,
    # upr, 0), sr = max(upl, upr, 0); with sl < 0 < sr the guard term is 0 and the fluxes are
    # HLL of the Burgers flux: velocity n carries F = 1/2 u_n u_p with the 1/2 on the whole
    # expression (the diffusive term included), a passive scalar q carries F = q u_p.
    a = straddling_states(2000, 11)
    upl, upr = a[:, 6], a[:, 7]
    sl, sr = np.minimum(upl, upr), np.maximum(upl, upr)
    assert np.all(sl < 0) and np.all(sr > 0)
    f = hll(a)
    np.testing.assert_array_equal(f[:, 0], sl)
    np.testing.assert_array_equal(f[:, 1], sr)
    for n in range(3):
        ul, ur = a[:, 2 * n], a[:, 2 * n + 1]
        exp = 0.5 * textbook_hll(ul * upl, ur * upr, ul, ur, sl, sr)
        # the terms are O(1): fluxes that cancel to ~0 carry ~1e-15 of rounding either way
        np.testing.assert_allclose(f[:, 2 + n], exp, rtol=1e-13, atol=1e-14)
    q = np.random.default_rng(12).uniform(-3, 3, size=(2000, 2))
    g = qflux(np.column_stack([upl, upr, q]))
    exp = textbook_hll(q[:, 0] * upl, q[:, 1] * upr, q[:, 0], q[:, 1], sl, sr)
    np.testing.assert_allclose(g, exp, rtol=1e-13, atol=1e-14)


def test_hll_one_signed_speeds_follow_the_guard():
    # With both normal velocities of one sign, sl * sr == 0 and the reference's denominator is
    # sr - sl + 1: the velocity flux is 1/2 sr u_n(L) upl / (sr + 1) for upl, upr >= 0 and
    # -1/2 sl u_n(R) upr / (1 - sl) for upl, upr <= 0; scalars the same without the 1/2.
    # Hand-worked cases (exact in binary):
    #   upl = 2, upr = 1:  sl = 0, sr = 2, u_x(L) = 3 -> 1/2 * 2 * 3 * 2 / 3 = 2; q(L) = 3 -> 4
    #   upl = -1, upr = -3: sl = -3, sr = 0, u_x(R) = 5 -> -1/2 * -3 * 5 * -3 / 4 = -45/8;
    #                       q(R) = 5 -> -45/4
    #   upl = upr = 0: sl = sr = 0 -> every flux 0
    #   upl = 0, upr = 1: sl = 0, sr = 1 -> upwind value 1/2 * 1 * u(L) * 0 / 2 = 0
    cases = np.array([[3, 7, 1, 1, 1, 1, 2, 1],
                      [7, 5, 1, 1, 1, 1, -1, -3],
                      [3, 7, 1, 1, 1, 1, 0, 0],
                      [3, 7, 1, 1, 1, 1, 0, 1]], dtype=np.float64)
    f = hll(cases)
    assert f[0, 2] == 2.0 and f[1, 2] == -45.0 / 8.0 and f[2, 2] == 0.0 and f[3, 2] == 0.0
    assert f[0, 3] == 0.5 * 2 * 1 * 2 / 3
    g = qflux(np.array([[2, 1, 3, 7], [-1, -3, 7, 5], [0, 0, 3, 7], [0, 1, 3, 7]], dtype=float))
    assert g[0] == 4.0 and g[1] == -45.0 / 4.0 and g[2] == 0.0 and g[3] == 0.0
    # random one-signed states against the closed forms
    r = np.random.default_rng(13)
    a = r.uniform(-2, 2, size=(1000, 8))
    a[:500, 6:8] = r.uniform(0, 2, size=(500, 2))
    a[500:, 6:8] = -r.uniform(0, 2, size=(500, 2))
    upl, upr = a[:, 6], a[:, 7]
    sl, sr = np.minimum(np.minimum(upl, upr), 0), np.maximum(np.maximum(upl, upr), 0)
    f = hll(a)
    pos = np.arange(1000) < 500
    for n in range(3):
        ul, ur = a[:, 2 * n], a[:, 2 * n + 1]
        exp = np.where(pos, 0.5 * sr * ul * upl / (sr + 1), -0.5 * sl * ur * upr / (1 - sl))
        np.testing.assert_allclose(f[:, 2 + n], exp, rtol=1e-13, atol=1e-15)
    q = r.uniform(-3, 3, size=(1000, 2))
    g = qflux(np.column_stack([upl, upr, q]))
    exp = np.where(pos, sr * upl * q[:, 0] / (sr + 1), -sl * upr * q[:, 1] / (1 - sl))
    np.testing.assert_allclose(g, exp, rtol=1e-13, atol=1e-15)


def test_calculate_fluxes_on_a_uniform_state_is_the_guarded_flux():
    # CalculateFluxes over whole blocks: with a uniform state every face has ql = qr = the
    # state, so every face flux is the one-signed closed form of its direction's velocity,
    # F_n = 1/2 |u_d| u_n u_d / (|u_d| + 1) (velocities), |u_d| q u_d / (|u_d| + 1) (scalars)
    # for u_d > 0, and -u_d ... / (1 - u_d) for u_d < 0; stored at every face of the block.
    m, b = burgers(nx=(16, 16, 16), block=(8, 8, 8), ns=2)
    U = b.get(0).reshape(m.field_shape(b.V))
    state = (0.75, -0.5, 0.25, 2.0, -1.5)
    for c in range(b.V):
        U[:, c] = state[c]
    b.set(0, U)
    lib().orc_burgers_fluxes(b.h)
    g, n = m.ng[0], m.n[0]
    for d in range(3):
        F = b.get(4 + d).reshape(U.shape)
        ud = state[d]
        s = abs(ud) / (abs(ud) + 1.0)
        hi = [g + n] * 3
        hi[d] += 1  # faces lo .. hi of the flux direction
        sub = F[:, :, g:hi[2], g:hi[1], g:hi[0]]
        for c in range(b.V):
            exp = (0.5 if c < 3 else 1.0) * s * state[c] * ud
            np.testing.assert_allclose(sub[:, c], exp, rtol=1e-15, atol=0)


def octant_boxes(lo=-0.5, hi=0.5):
    mid = 0.5 * (lo + hi)
    out = []
    for s1 in (0, 1):           # the reference's octant order: x outer, then y, then z
        for s2 in (0, 1):
            for s3 in (0, 1):
                out.append(([mid if s else lo for s in (s1, s2, s3)],
                            [hi if s else mid for s in (s1, s2, s3)]))
    return out


def octant_constant_field(m, V, coef):
    """U = coef[o, v] on every cell whose centre lies in octant o (octant 4 s1 + 2 s2 + s3)."""
    U = np.zeros(m.field_shape(V))
    g, n = m.ng[0], m.n[0]
    for gid in range(m.nblocks):
        bk = m.block(gid)
        xc = [bk["xmin"][d] + (np.arange(-g, n + g) + 0.5) * (bk["xmax"][d] - bk["xmin"][d]) / n
              for d in range(3)]
        s = [(x > 0).astype(int) for x in xc]
        o = 4 * s[0][None, None, :] + 2 * s[1][None, :, None] + s[2][:, None, None]
        for v in range(V):
            U[gid, v] = coef[o, v]
    return U


@pytest.mark.parametrize("refined", [False, True])
def test_mass_history_octant_constant_fields(refined):
    # burgers_package.cpp:96-135, 406-439: mass_o = sum over cells of mask_o(x_c) U^2 vol /
    # mesh_vol. With U = c[o, v] on octant o every octant holds 1/8 of the unit mesh volume at
    # any refinement, so mass_o = sum_v c[o, v]^2 / 8, in the order x outer, then y, then z.
    boxes = [((-0.1, 0.1, -0.1, 0.1, -0.1, 0.1), 1)] if refined else []
    m = OracleMesh((32, 32, 32), (8, 8, 8), 4, boxes=boxes)
    b = OracleBurgers(m, 2, 0.4, 0)
    coef = np.array([[1, 2, 3, 4, 5], [-1, 0.5, 0, 2, 1], [3, 1, 1, 1, 1], [0, 0, 0, 0, 4],
                     [2, 2, 2, 2, 2], [0.25, -3, 1, 0, 1], [1, 1, -1, -1, 1], [5, 0, 0, 0, 0]],
                    dtype=np.float64)
    b.set(0, octant_constant_field(m, b.V, coef))
    got = [lib().orc_burgers_mass_history(b.h, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), 1.0)
           for lo, hi in octant_boxes()]
    exp = (coef ** 2).sum(axis=1) / 8.0
    np.testing.assert_allclose(got, exp, rtol=1e-13)
    # the mesh volume divides: a domain volume of 2 halves every mass
    half = lib().orc_burgers_mass_history(b.h, (C.c_double * 3)(*octant_boxes()[0][0]),
                                          (C.c_double * 3)(*octant_boxes()[0][1]), 2.0)
    assert abs(half - exp[0] / 2) < 1e-13 * exp[0]


def uniform_run(cfl=0.375, **limits):
    """Uniform state (0.75, -0.5, 0.25) on 32^3 cells of width 1/32: the CFL estimate is
    cfl / (1.5 * 32) = 2^-7 exactly for cfl = 0.375, and the state is steady."""
    m, b = burgers(cfl=cfl)
    U = b.get(0).reshape(m.field_shape(b.V))
    vel = (0.75, -0.5, 0.25)
    for c in range(b.V):
        U[:, c] = vel[c] if c < 3 else 2.0
    b.set(0, U)
    b.set_time_limits(**limits)
    lib().orc_burgers_initialize(b.h)
    return b, U


def test_time_limits_tlim_cuts_the_last_cycle_and_stops():
    # SetGlobalTimeStep's end clamp (driver.cpp:267-268) and KeepGoing (basic_types.hpp:241):
    # with dt = 2^-7 and tlim = 2.5 * 2^-7 the third cycle is 2^-8 long, ends on tlim exactly,
    # and the loop stops; the dt computed after it is the unclamped 2^-7
    b, U = uniform_run(tlim=2.5 * 2.0 ** -7)
    assert b.dt == 2.0 ** -7
    assert b.cycles(10) == 3
    assert b.time == 2.5 * 2.0 ** -7 and b.ncycle == 3 and b.dt == 2.0 ** -7
    assert b.cycles(1) == 0
    assert np.array_equal(b.get(0).reshape(U.shape), U)


def test_time_limits_nlim_stops():
    b, _ = uniform_run(nlim=2)
    assert b.cycles(10) == 2 and b.time == 2.0 ** -6


@pytest.mark.parametrize("limits,dts", [
    (dict(dt_user=2.0 ** -9), [2.0 ** -9] * 4),
    (dict(dt_force=2.0 ** -8), [2.0 ** -8] * 4),
    # dt_init bounds only the first dt; then dt doubles until the estimate binds
    (dict(dt_init=2.0 ** -10), [2.0 ** -10, 2.0 ** -9, 2.0 ** -8, 2.0 ** -7, 2.0 ** -7]),
    # forced, dt_init may exceed the estimate on the first cycle only
    (dict(dt_init=2.0 ** -5, dt_init_force=True), [2.0 ** -5, 2.0 ** -7, 2.0 ** -7]),
    (dict(dt_ceil=2.0 ** -8), [2.0 ** -8] * 3),
    (dict(dt_floor=2.0 ** -6), [2.0 ** -6] * 3),
], ids=["dt_user", "dt_force", "dt_init", "dt_init_force", "dt_ceil", "dt_floor"])
def test_time_limits_dt_bounds(limits, dts):
    # SetGlobalTimeStep (driver.cpp:210-235) with each <parthenon/time> bound against the
    # estimate 2^-7: the dt of every cycle and the time they add up to
    b, _ = uniform_run(**limits)
    t = 0.0
    for dt in dts:
        assert b.dt == dt, (limits, b.ncycle, b.dt, dt)
        assert b.cycles(1) == 1
        t += dt
        assert b.time == t
