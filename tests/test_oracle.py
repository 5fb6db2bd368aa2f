"""Pins of the CPU oracle (no GPU).

The reference ships no tests, fixtures or golden vectors, and it cannot be
built here (nvcc and the unvendored gxl_lib are absent), so each oracle
function is pinned by closed forms, invariants and brute force chosen so
that a dropped term, a wrong sign or index, or a transposed operand fails
at least one check:

  NASA-7 thermo (Thermo.cu:7-63)      dh/dT = cp by finite differences;
                                       constant-cp closed form; linear
                                       extension and frozen cp below t_low
  Newton T inversion (Thermo.cu:89-125)   round trip T -> e(T) -> T; constant-cp
                                       closed form T = (e - sum Y r a5)/(sum Y r (a0-1))
  Sutherland (Transport.cu:7-10)        closed form
  Wilke/Eucken (Transport.cu:43-96)      one species = Chapman-Enskog closed
                                       form; N identical species = one;
                                       long-double brute force; rho*D closed form
  limiters (Limiter.cuh:13-16)           closed forms, symmetry
  MUSCL + AUSM+ (InviscidScheme.cu)    F(U,U) = physical flux |k|; supersonic
                                       upwinding F = F(U_L); linear data;
                                       first order takes the cell values
  metrics (Mesh.cpp:226-488)           Cartesian closed form; 2-D discrete GCL
  full step                            freestream preservation on a curved
                                       2-D grid; mass conservation in a closed
                                       slip-wall box; mirror symmetry of the
                                       slip wall; i<->j transpose symmetry;
                                       2-block split = 1 block; RK3 order
  residual (Driver.cu:219-277)         numpy brute force
"""
import math

import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases

RU = 8314.462618


@pytest.fixture(scope="module")
def orc(oracle_lib):
    return oracle_lib


@pytest.fixture(scope="module")
def sp9():
    return cases.synthetic_species(9, 2024)


def one_species(a0=3.5, a5=-1000.0, w=28.0, eps=100.0, sigma=3.5, t_low=200.0, t_mid=1000.0):
    c = np.zeros((1, 7))
    c[0, 0] = a0
    c[0, 5] = a5
    return {"mw": np.array([w]), "t_low": np.array([t_low]), "t_mid": np.array([t_mid]),
            "t_high": np.array([3000.0]), "coeff_lo": c, "coeff_hi": c.copy(), "eps_k": np.array([eps]),
            "sigma": np.array([sigma])}


# ---------------------------------------------------------------- thermo ---
def test_enthalpy_derivative_is_cp(orc, sp9):
    for t in (250.0, 700.0, 999.0, 1500.0, 2900.0):
        hp, _ = orc.enthalpy_cp(sp9, t + 1e-3)
        hm, _ = orc.enthalpy_cp(sp9, t - 1e-3)
        _, cp = orc.enthalpy_cp(sp9, t)
        np.testing.assert_allclose((hp - hm) / 2e-3, cp, rtol=1e-7)


def test_enthalpy_below_t_low(orc, sp9):
    h0, cp0 = orc.enthalpy_cp(sp9, 200.0)
    h, cp = orc.enthalpy_cp(sp9, 150.0)
    np.testing.assert_allclose(cp, cp0, rtol=1e-15)            # frozen cp
    np.testing.assert_allclose(h, h0 + cp0 * (150.0 - 200.0), rtol=1e-13)  # linear extension


def test_constant_cp_closed_form(orc):
    s = one_species(a0=3.5, a5=-1000.0, w=28.0)
    h, cp = orc.enthalpy_cp(s, 800.0)
    r = RU / 28.0
    assert abs(cp[0] - 3.5 * r) <= 1e-13 * cp[0]
    assert abs(h[0] - r * (3.5 * 800.0 - 1000.0)) <= 1e-13 * abs(h[0])


def test_newton_round_trip(orc, sp9):
    rng = np.random.default_rng(1)
    for _ in range(20):
        Y = rng.dirichlet(np.ones(9))
        T = rng.uniform(300, 2500)
        h, _ = orc.enthalpy_cp(sp9, T)
        r_mix = RU * np.sum(Y / sp9["mw"])
        e = np.sum(Y * h) - r_mix * T
        t, it = orc.newton_temperature(sp9, Y, e, T * 1.3)
        assert abs(t - T) <= 1e-9 and it < 50


def test_newton_constant_cp_closed_form(orc):
    s = one_species(a0=3.5, a5=-1000.0, w=28.0)
    r = RU / 28.0
    e = 5.0e5
    t, _ = orc.newton_temperature(s, np.array([1.0]), e, 400.0)
    exact = (e - r * (-1000.0)) / (r * (3.5 - 1.0))
    assert abs(t - exact) <= 1e-9


def test_newton_reference_stop_rule(orc, sp9):
    """With the reference's |1-T1/T|<=1e-3 rule the result is within the
    quadratic-convergence bound of the exact T."""
    Y = np.full(9, 1 / 9)
    T = 1234.0
    h, _ = orc.enthalpy_cp(sp9, T)
    e = np.sum(Y * h) - RU * np.sum(Y / sp9["mw"]) * T
    t, _ = orc.newton_temperature(sp9, Y, e, 1200.0, quirks=abi.Q_NEWTON_REL)
    assert abs(t - T) < 1e-2


# ------------------------------------------------------------- transport ---
def test_sutherland(orc):
    for t in (200.0, 273.0, 300.0, 1000.0):
        assert abs(orc.sutherland(t) - 1.716e-5 * (t / 273) ** 1.5 * 384 / (t + 111)) <= 1e-18


def chapman_enskog(t, w, eps, sigma):
    tdl = t / eps
    omega = 1.147 * tdl ** -0.145 + (tdl + 0.5) ** -2
    return 2.6693e-6 * math.sqrt(w) / sigma ** 2 * math.sqrt(t) / omega


def test_single_species_transport(orc):
    s = one_species(a0=3.5, w=28.0, eps=97.5, sigma=3.62)
    mu, kap, rd = orc.transport(s, 1200.0, np.array([1.0]), schmidt=0.9)
    mu1 = chapman_enskog(1200.0, 28.0, 97.5, 3.62)
    cp = 3.5 * RU / 28.0
    assert abs(mu - mu1) <= 1e-14 * mu1
    assert abs(kap - mu1 * (cp + 1.25 * RU / 28.0)) <= 1e-13 * kap  # 1.065-0.065 = 1
    assert abs(rd[0] - mu / 0.9) <= 1e-14 * rd[0]                   # X = 1 branch


def test_identical_species(orc):
    one = one_species(w=16.0, eps=150.0, sigma=3.0)
    many = {k: (np.repeat(v, 4, axis=0) if v.ndim else v) for k, v in one.items()}
    Y = np.array([0.1, 0.2, 0.3, 0.4])
    mu1, k1, _ = orc.transport(one, 900.0, np.array([1.0]))
    mu4, k4, rd = orc.transport(many, 900.0, Y)
    assert abs(mu4 - mu1) <= 1e-14 * mu1
    # phi = 1, so the Mason-Saxena denominator is 1.065 - 0.065 X_m
    k_exp = k1 * np.sum(Y / (1.065 - 0.065 * Y))
    assert abs(k4 - k_exp) <= 1e-13 * k_exp
    # rho*D = (1-Y) mu / ((1-X) Sc) with X = Y for equal weights -> mu/Sc
    np.testing.assert_allclose(rd, mu4 / 0.9, rtol=1e-14)


def test_wilke_long_double_brute_force(orc, sp9):
    rng = np.random.default_rng(5)
    ld = np.longdouble
    for _ in range(5):
        Y = rng.dirichlet(np.ones(9))
        T = rng.uniform(300, 2500)
        mu, kap, rd = orc.transport(sp9, T, Y)
        W = sp9["mw"].astype(ld)
        Yl = Y.astype(ld)
        Tl = ld(T)
        mw_mix = 1 / np.sum(Yl / W)
        X = Yl * mw_mix / W
        tdl = Tl / sp9["eps_k"].astype(ld)
        om = ld(1.147) * tdl ** ld(-0.145) + (tdl + ld(0.5)) ** -2
        vis = ld(2.6693e-6) * np.sqrt(W) / sp9["sigma"].astype(ld) ** 2 * np.sqrt(Tl) / om
        _, cp = orc.enthalpy_cp(sp9, min(T, 3000.0))
        lam = vis * (cp.astype(ld) + ld(1.25) * ld(RU) / W)
        phi = (1 + np.sqrt(vis[:, None] / vis[None, :]) * (W[None, :] / W[:, None]) ** ld(0.25)) ** 2 / np.sqrt(
            8 * (1 + W[:, None] / W[None, :]))
        s = phi @ X
        mu_b = np.sum(vis * X / s)
        k_b = np.sum(lam * X / (ld(1.065) * s - ld(0.065) * X))
        assert abs(float(mu_b) - mu) <= 1e-13 * mu
        assert abs(float(k_b) - kap) <= 1e-13 * kap
        rd_b = (1 - Yl) * mu_b / ((1 - X) * ld(0.9))
        np.testing.assert_allclose(rd, rd_b.astype(float), rtol=1e-13)


# -------------------------------------------------------------- limiters ---
def test_limiters(orc):
    assert orc.limiter(0, 1.0, 3.0) == 1.0
    assert orc.limiter(0, -2.0, -0.5) == -0.5
    assert orc.limiter(0, 1.0, -1.0) == 0.0
    assert orc.limiter(0, 0.0, 5.0) == 0.0
    assert orc.limiter(1, 2.0, 2.0) == pytest.approx(2.0, rel=1e-15)
    assert orc.limiter(1, 1.0, -1.0) == 0.0
    assert orc.limiter(1, 1.0, 3.0) == pytest.approx(1 * 3 * 4 / 10, rel=1e-15)
    rng = np.random.default_rng(2)
    for a, b in rng.normal(size=(50, 2)):
        for tag in (0, 1):
            assert orc.limiter(tag, a, b) == orc.limiter(tag, b, a)  # bitwise symmetric


# --------------------------------------------------------------- AUSM+ -----
def physical_flux(q, k, sp, gamma=1.4):
    rho, u, v, w, p = q[:5]
    kk = np.asarray(k)
    un = kk @ np.array([u, v, w])
    if sp is None:
        E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
        Y = np.zeros(0)
    else:
        Y = q[5:]
        T = p / (rho * RU * np.sum(Y / sp["mw"]))
        from oracle import pyoracle
        h, _ = pyoracle.enthalpy_cp(sp, T)
        E = rho * (0.5 * (u * u + v * v + w * w) + np.sum(h * Y)) - p
    f = [rho * un, rho * u * un + p * kk[0], rho * v * un + p * kk[1], rho * w * un + p * kk[2], (E + p) * un]
    return np.array(f + list(rho * Y * un))


@pytest.mark.parametrize("mach", [0.3, -0.7, 1.8, -2.5])
def test_ausm_consistency(orc, sp9, mach):
    for sp in (None, sp9):
        Y = [] if sp is None else list(np.random.default_rng(3).dirichlet(np.ones(9)))
        rho, p = 1.1, 9.0e4
        if sp is None:
            c = math.sqrt(1.4 * p / rho)
        else:
            from oracle import pyoracle
            T = p / (rho * RU * np.sum(np.array(Y) / sp["mw"]))
            _, cpi = pyoracle.enthalpy_cp(sp, T)
            cpm = np.sum(cpi * Y)
            g = cpm / (cpm - RU * np.sum(np.array(Y) / sp["mw"]))
            c = math.sqrt(g * p / rho)
        metric = np.array([[0.8, 0.3, -0.1], [0.8, 0.3, -0.1]])
        jac = np.array([2.0, 2.0])
        k = 0.5 * (jac[0] * metric[0] + jac[1] * metric[1])
        nrm = np.linalg.norm(k)
        vel = mach * c * k / nrm + np.array([-0.3, 0.8, 0.0]) * 17.0  # add a tangential part
        vel -= (vel @ k / nrm - mach * c) * k / nrm
        q = np.array([rho, *vel, p] + Y)
        pv = np.tile(q, (4, 1))
        f = orc.inviscid_face(sp, {"reconstruction": 2, "limiter": 1, "r_air": 287.0}, pv, metric, jac)
        np.testing.assert_allclose(f, physical_flux(q, k, sp), rtol=1e-12, atol=1e-12 * abs(p) * nrm)


def test_ausm_supersonic_upwind(orc):
    ql = np.array([1.0, 900.0, 10.0, 0.0, 1.0e5])
    qr = np.array([0.8, 800.0, -5.0, 3.0, 0.7e5])
    pv = np.array([ql, ql, qr, qr])  # first order at the face: L = ql, R = qr
    metric = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    jac = np.array([1.0, 1.0])
    f = orc.inviscid_face(None, {"reconstruction": 1, "limiter": 0, "r_air": 287.0}, pv, metric, jac)
    np.testing.assert_allclose(f, physical_flux(ql, [1.0, 0.0, 0.0], None), rtol=1e-14)


def test_first_order_takes_the_cell_values(orc):
    """First-order reconstruction (Reconstruction.cu:12-20): the face states
    are the two adjacent cell values whatever the outer stencil points hold,
    so the flux equals the MUSCL flux of data that is flat on each side."""
    ql = np.array([1.0, 300.0, 20.0, -5.0, 1.0e5])
    qr = np.array([0.9, 250.0, -10.0, 4.0, 0.9e5])
    far_l = np.array([1.3, 100.0, 0.0, 0.0, 1.4e5])
    far_r = np.array([0.5, 600.0, 30.0, 9.0, 0.5e5])
    metric = np.array([[0.8, 0.3, 0.1], [0.7, 0.2, 0.2]])
    jac = np.array([1.1, 0.9])
    f1 = orc.inviscid_face(None, {"reconstruction": 1, "limiter": 1, "r_air": 287.0},
                           np.array([far_l, ql, qr, far_r]), metric, jac)
    f2 = orc.inviscid_face(None, {"reconstruction": 2, "limiter": 1, "r_air": 287.0},
                           np.array([ql, ql, qr, qr]), metric, jac)
    assert np.array_equal(f1, f2)
    f3 = orc.inviscid_face(None, {"reconstruction": 2, "limiter": 1, "r_air": 287.0},
                           np.array([far_l, ql, qr, far_r]), metric, jac)
    assert not np.allclose(f1, f3)


def test_muscl_reproduces_linear_data(orc):
    """With linear data the limited slopes equal the data slope, so the face
    states are the linear interpolant and left == right: AUSM+ then gives
    the physical flux of the midpoint state."""
    base = np.array([1.0, 300.0, 20.0, -5.0, 1.0e5])
    d = np.array([0.01, 2.0, 0.5, 0.1, 500.0])
    pv = np.array([base + (i - 1) * d for i in range(4)])
    mid = base + 0.5 * d
    metric = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    jac = np.array([1.0, 1.0])
    for tag in (0, 1):
        f = orc.inviscid_face(None, {"reconstruction": 2, "limiter": tag, "r_air": 287.0}, pv, metric, jac)
        np.testing.assert_allclose(f, physical_flux(mid, [1, 0, 0], None), rtol=1e-12)


# -------------------------------------------------------------- metrics ----
def test_cartesian_metrics(orc):
    from oracle.pyoracle import Oracle
    mx, my, mz = 8, 7, 6
    dx, dy, dz = 1e-3, 2e-3, 0.5e-3
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    c = cases.c1(mx, my, mz, n_spec=2)
    b = c.blocks[0]
    b.x, b.y, b.z = [np.ascontiguousarray(a, dtype=float).ravel() for a in (i * dx, j * dy, k * dz)]
    o = Oracle(c)
    m = o.download(0, abi.METRIC)
    np.testing.assert_allclose(m[0], 1 / dx, rtol=1e-12)
    np.testing.assert_allclose(m[4], 1 / dy, rtol=1e-12)
    np.testing.assert_allclose(m[8], 1 / dz, rtol=1e-12)
    for idx in (1, 2, 3, 5, 6, 7):  # corner ghosts carry ulp-level round-off
        assert np.abs(m[idx]).max() <= 1e-12 / dx
    np.testing.assert_allclose(m[9], dx * dy * dz, rtol=1e-12)


def run(case, steps=0, do_init=True):
    from oracle.pyoracle import Oracle
    o = Oracle(case)
    o.upload_initial()
    if do_init:
        o.init_conservative()
    if steps:
        o.step(steps)
    return o


def uniform_field(case, rho=1.2, u=(120.0, -40.0, 0.0), p=9.0e4):
    for b in range(len(case.blocks)):
        bv = case.bv0[b]
        bv[0] = rho
        bv[1], bv[2], bv[3] = u
        bv[4] = p
        bv[5] = p / (rho * case.params.get("r_air", 287.0))
    return case


def test_freestream_preservation_curved_2d(orc):
    """Uniform flow on a curved 2-D grid: the discrete GCL of the cross-product
    metrics holds exactly in 2-D, so dq vanishes to round-off."""
    c = cases.c0(40, 30)
    b = c.blocks[0]
    x, y, _ = cases.cartesian_grid(40, 30, 1, 1e-3, perturb=0.4, seed=3)
    b.x, b.y = x, y
    c.bc = []
    b.bc = [((0, 0, 0), (0, 29, 0), abi.BC_OUTFLOW, 0), ((39, 0, 0), (39, 29, 0), abi.BC_OUTFLOW, 0),
            ((0, 0, 0), (39, 0, 0), abi.BC_OUTFLOW, 0), ((0, 29, 0), (39, 29, 0), abi.BC_OUTFLOW, 0)]
    c.inflow = []
    uniform_field(c)
    o = run(c)
    o.rhs(-1)
    dq = o.download(0, abi.DQ)
    m = o.download(0, abi.METRIC)
    scale = 9.0e4 * np.abs(m[9]).max() * np.abs(m[0]).max()  # p * |k|
    assert np.abs(dq[:, :, 1:-1, 1:-1]).max() <= 1e-12 * scale


def closed_box(nx=24, ny=20):
    c = cases.c0(nx, ny)
    b = c.blocks[0]
    X, Y = nx - 1, ny - 1
    b.bc = [((0, 0, 0), (0, Y, 0), abi.BC_SYMMETRY, 0), ((X, 0, 0), (X, Y, 0), abi.BC_SYMMETRY, 0),
            ((0, 0, 0), (X, 0, 0), abi.BC_SYMMETRY, 0), ((0, Y, 0), (X, Y, 0), abi.BC_SYMMETRY, 0)]
    c.inflow = []
    bv = c.bv0[0]
    bv[1] = 30.0 * np.sin(np.linspace(0, 3, bv.shape[-1]))[None, None, :]
    bv[2] = -20.0 * np.cos(np.linspace(0, 2, bv.shape[-2]))[None, :, None]
    return c


def test_mass_conservation_closed_box(orc):
    """Slip walls on all sides.  A wall vertex's control volume straddles the
    wall and the mirrored ghosts make the data symmetric about it, so the
    unfolded (mirrored) domain is closed: sum(w rho J) is constant with
    w = 1/2 per wall the vertex lies on (1/4 at corners)."""
    c = closed_box()
    o = run(c)
    J = o.download(0, abi.METRIC)[9][2, 2:-2, 2:-2]
    w = np.ones_like(J)
    w[0, :] *= 0.5
    w[-1, :] *= 0.5
    w[:, 0] *= 0.5
    w[:, -1] *= 0.5
    m0 = np.sum(o.download(0, abi.CV)[0][2, 2:-2, 2:-2] * J * w)
    o.step(5)
    m1 = np.sum(o.download(0, abi.CV)[0][2, 2:-2, 2:-2] * J * w)
    assert abs(m1 - m0) <= 1e-13 * m0


def test_slip_wall_mirror_symmetry(orc):
    """A mirror-symmetric flow on [-L, L] equals the half domain [0, L] with a
    slip wall at y = 0."""
    nx, ny = 24, 15
    full = cases.c0(nx, 2 * ny - 1, seed=1)
    half = cases.c0(nx, ny, seed=1)
    jj = np.arange(-2, 2 * ny - 1 + 2) - (ny - 1)
    jh = np.arange(-2, ny + 2)
    fb = full.bv0[0]
    fb[0] = 1.17 * (1 + 0.05 * np.exp(-(jj[None, :, None] / 4.0) ** 2) * np.cos(np.arange(-2, nx + 2) / 5.0)[None, None, :])
    fb[2] = 7.0 * np.sign(jj)[None, :, None] * np.sin(np.abs(jj) / 3.0)[None, :, None]
    fb[5] = fb[4] / (fb[0] * 287.0)
    hb = half.bv0[0]
    sel = jh + (ny - 1) + 2  # padded index of full-domain row j + ny - 1
    hb[:] = fb[:, :, np.clip(sel, 0, fb.shape[2] - 1), :]
    half.blocks[0].bc[2] = ((0, 0, 0), (nx - 1, 0, 0), abi.BC_SYMMETRY, 0)
    of = run(full, 3)
    oh = run(half, 3)
    a = of.download(0, abi.BV)[:, 2, 2 + ny - 1:2 + 2 * ny - 1, 2:-2]
    b = oh.download(0, abi.BV)[:, 2, 2:2 + ny, 2:-2]
    np.testing.assert_allclose(b[[0, 1, 4, 5]], a[[0, 1, 4, 5]], rtol=1e-11)
    assert np.abs(b[2] - a[2]).max() <= 1e-9 * np.abs(a[1]).max()


def test_transpose_symmetry(orc):
    """Swapping i and j of a cubic-cell 3-D case (corrected quirks, which are
    direction-free) transposes the result; compared away from the block
    edges, whose corner ghosts depend on the boundary order."""
    c = cases.c1(12, 12, 8, n_spec=2, quirks=0)
    t = cases.c1(12, 12, 8, n_spec=2, quirks=0)
    bv = c.bv0[0]
    bt = bv.transpose(0, 1, 3, 2).copy()
    bt[[1, 2]] = bt[[2, 1]]
    t.bv0[0] = np.ascontiguousarray(bt)
    t.yk0[0] = np.ascontiguousarray(c.yk0[0].transpose(0, 1, 3, 2))
    oc = run(c, 1)
    ot = run(t, 1)
    a = oc.download(0, abi.CV)
    b = ot.download(0, abi.CV).transpose(0, 1, 3, 2)
    b[[1, 2]] = b[[2, 1]]
    sl = np.s_[:, 4:-4, 4:-4, 4:-4]
    scale = np.abs(a[sl]).reshape(a.shape[0], -1).max(axis=1)
    scale[1:4] = scale[1:4].max()
    err = np.abs(a[sl] - b[sl]).reshape(a.shape[0], -1).max(axis=1) / scale
    assert err.max() <= 1e-13


def test_two_block_split_equals_one_block(orc):
    c = cases.c1(16, 12, 10)
    s = cases.split_i(c, 9)
    o1 = run(c, 3)
    o2 = run(s, 3)
    a = o1.download(0, abi.CV)
    A, B = o2.download(0, abi.CV), o2.download(1, abi.CV)
    ia, ib = a[..., 2:-2, 2:-2, 2:2 + 9], a[..., 2:-2, 2:-2, 2 + 8:-2]
    assert np.abs(A[..., 2:-2, 2:-2, 2:-2] - ia).max() <= 1e-13 * np.abs(ia).max()
    assert np.abs(B[..., 2:-2, 2:-2, 2:-2] - ib).max() <= 1e-13 * np.abs(ib).max()


def test_rk3_third_order(orc):
    """Fixed-dt SSP-RK3: halving dt cuts the error against a fine-dt solution
    by 2^3.  First-order reconstruction keeps the semi-discrete operator
    smooth (the flow is supersonic, so AUSM+ is F(U_L)); with a limiter the
    clipping at the inflow junction masks the temporal order."""
    def solve(dt, n):
        c = cases.c0(32, 16, cfl=0.5)
        c.params["reconstruction"] = 1
        o = run(c)
        for _ in range(n):
            o.store_last_step()
            for s in range(3):
                o.rhs(-1)
                o.set_dt(dt)
                o.update(s)
                o.apply_bc()
                o.eliminate_k_gradient()
                o.update_properties()
        return o.download(0, abi.CV)[:, 2, 2:-2, 2:-2]
    T = 8e-7
    ref = solve(T / 64, 64)
    e1 = np.abs(solve(T / 8, 8) - ref)[0].max()
    e2 = np.abs(solve(T / 16, 16) - ref)[0].max()
    assert 7.0 < e1 / e2 < 9.5, (e1, e2)


def test_residual_brute_force(orc):
    c = cases.c1(10, 9, 8, n_spec=2)
    o = run(c)
    bv0 = o.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2]
    o.step(1)
    bv1 = o.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2]
    r = o.residual()

    def q(bv):
        return [bv[0], np.sqrt(bv[1] ** 2 + bv[2] ** 2 + bv[3] ** 2), bv[4], bv[5]]
    exp = [math.sqrt(np.mean((a - b) ** 2)) for a, b in zip(q(bv1), q(bv0))]
    np.testing.assert_allclose(r, exp, rtol=1e-12)


def test_inflow_state(orc):
    """Inflow (Mach 2, p, T) on the air path: rho = p/(RT), u = 2 sqrt(gamma R T)
    (BoundCond.cu:68-103), written into the ghosts by apply_inflow."""
    c = cases.c0(16, 12)
    o = run(c)
    bv = o.download(0, abi.BV)
    g = bv[:, 2, 2:-2, 0:2]
    np.testing.assert_allclose(g[0], 101325 / (287 * 300), rtol=1e-15)
    np.testing.assert_allclose(g[1], 2 * math.sqrt(1.4 * 287 * 300), rtol=1e-15)
    assert np.all(g[2] == 0) and np.all(g[4] == 101325) and np.all(g[5] == 300)
