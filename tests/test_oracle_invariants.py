"""Pins of the oracle that do not need the reference binary: invariants,
closed forms, hand-derived stale values of the staged step, and analytic
flows.  Each check is chosen so that a dropped term, a wrong sign or index,
or a transposed operand in the oracle fails it."""
import math

import numpy as np
import pytest

from oracle.pyoracle import (Boundary, NOSLIP, PERIODIC, UNSET, UP, DOWN, EAST, WEST, NORTH,
                             SOUTH, channel_bc)

INTERIOR = (slice(1, -1), slice(1, -1), slice(1, -1))


def feq_np(c, w, rho, u):
    cu = c @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))


# --------------------------------------------------------------- collide --
@pytest.mark.parametrize("q", [9, 19, 27])
def test_collide_feq_fixed_point(orc, q):
    """With F = 0, feq(rho, u) is a fixed point of BGK."""
    X, Y, Z = 3, 1 if q == 9 else 3, 3
    st = orc.lit_state(q, X, Y, Z)
    rng = np.random.default_rng(q)
    for idx in np.ndindex(st["rho"].shape):
        rho = rng.uniform(0.9, 1.1)
        u = rng.uniform(-0.05, 0.05, 3)
        if q == 9:
            u[1] = 0.0
        st["rho"][idx] = rho
        st["u"][idx] = u
        st["n"][idx] = orc.feq(q, rho, u)
    before = st["n"].copy()
    orc.lit_collide(st, 0.83, (0.0, 0.0, 0.0))
    assert np.allclose(st["n"], before, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("q", [9, 19, 27])
def test_collide_conserves_mass_and_adds_force(orc, q):
    """Sum f is conserved; sum c f gains exactly F (omega * tau = 1)."""
    X, Y, Z = 4, 1 if q == 9 else 3, 3
    c, w, _, _ = orc.tables(q)
    st = orc.lit_state(q, X, Y, Z)
    rng = np.random.default_rng(10 + q)
    st["n"][...] = w * rng.uniform(0.8, 1.2, st["n"].shape)
    orc.lit_moments(st)  # rho, u consistent with n on the interior
    F = np.array([3e-3, -2e-3 if q != 9 else 0.0, 1e-3])
    n0 = st["n"][INTERIOR].copy()
    orc.lit_collide(st, 0.71, F)
    n1 = st["n"][INTERIOR]
    assert np.allclose(n1.sum(-1), n0.sum(-1), rtol=1e-14)
    dj = n1 @ c - n0 @ c
    assert np.allclose(dj, F, rtol=1e-10, atol=1e-15)


def test_collide_single_cell_brute_force(orc):
    """One D3Q19 cell against a direct evaluation of the BGK update."""
    q = 19
    c, w, _, _ = orc.tables(q)
    st = orc.lit_state(q, 1, 1, 1)
    rho, u = 1.037, np.array([0.021, -0.013, 0.007])
    n = w * (1 + 0.1 * np.sin(np.arange(q)))
    st["n"][1, 1, 1] = n
    st["rho"][1, 1, 1] = rho
    st["u"][1, 1, 1] = u
    tau, F = 0.66, np.array([1e-3, 2e-3, -3e-3])
    orc.lit_collide(st, tau, F)
    us = u + tau * F / rho
    exp = n + (feq_np(c.astype(float), w, rho, us) - n) / tau
    assert np.allclose(st["n"][1, 1, 1], exp, rtol=1e-14)


# ------------------------------------------------------------------ push --
@pytest.mark.parametrize("q", [9, 19, 27])
def test_push_moves_every_value_by_its_velocity(orc, q):
    X, Y, Z = 5, 1 if q == 9 else 4, 3
    c, _, _, _ = orc.tables(q)
    st = orc.lit_state(q, X, Y, Z)
    st["n"][...] = np.arange(st["n"].size, dtype=float).reshape(st["n"].shape) + 1.0
    src = st["n"].copy()
    orc.lit_push(st)  # st["n"] is now the pushed array
    out = st["n"]
    written = np.zeros(out.shape, bool)
    for k in range(q):
        cx, cy, cz = c[k]
        dst = out[1 + cz:Z + 1 + cz, 1 + cy:Y + 1 + cy, 1 + cx:X + 1 + cx, k]
        assert np.array_equal(dst, src[1:-1, 1:-1, 1:-1, k])
        written[1 + cz:Z + 1 + cz, 1 + cy:Y + 1 + cy, 1 + cx:X + 1 + cx, k] = True
    assert np.all(out[~written] == 0.0)  # ntmp started at zero (SimData.cpp:53-56)


# ------------------------------------------------------ no-slip, periodic --
def test_noslip_reflects_wall_bound_populations(orc):
    """Every fluid cell next to the down wall gets back, in direction kp, the
    value that arrived at the wall cell c = y - c_kp in direction rev(kp)."""
    q, X, Y, Z = 19, 5, 4, 5
    c, _, _, rev = orc.tables(q)
    bc = Boundary(type=[UNSET, UNSET, UNSET, UNSET, UNSET, NOSLIP], vel={DOWN: (0, 0, 0)})
    st = orc.lit_state(q, X, Y, Z)
    st["n"][...] = np.random.default_rng(0).random(st["n"].shape)
    before = st["n"].copy()
    assert orc.lit_noslip(st, bc) == 0
    for kp in range(1, q):
        if c[kp][2] != 1:
            continue
        for y in range(1, Y + 1):
            for x in range(1, X + 1):
                cx, cy = x - c[kp][0], y - c[kp][1]
                dest = (2, y, x, kp)
                if 1 <= cx <= X and 1 <= cy <= Y:
                    assert st["n"][dest] == before[1, cy, cx, rev[kp]]


def test_noslip_moving_wall_term(orc):
    """n(y,kp) = n(c,k) - 2 w_k rho(y) (c_k . u_b) / cs2 for a moving lid."""
    q, X, Y, Z = 19, 4, 4, 4
    c, w, _, rev = orc.tables(q)
    ub = np.array([0.1, 0.05, 0.0])
    bc = Boundary(type=[UNSET] * 4 + [NOSLIP, UNSET], vel={UP: tuple(ub)})
    st = orc.lit_state(q, X, Y, Z)
    rng = np.random.default_rng(3)
    st["n"][...] = rng.random(st["n"].shape)
    st["rho"][...] = rng.uniform(0.9, 1.1, st["rho"].shape)
    before = st["n"].copy()
    orc.lit_noslip(st, bc)
    kp = [k for k in range(q) if tuple(c[k]) == (1, 0, -1)][0]
    k = rev[kp]
    y = (Z - 1, 2, 3)
    cc = (Z, 2, 2)
    exp = before[cc + (k,)] - 2 * w[k] * st["rho"][y] * (c[k] @ ub) * 3.0
    assert st["n"][y + (kp,)] == pytest.approx(exp, rel=1e-15)


def test_noslip_in_plane_serial_order_d2q9(orc):
    """In-plane entries of the wall layer follow the serial x, kp loop: after
    the pass n(x,1) holds the old n(x-1,2) and n(x,2) is unchanged (u_b = 0)."""
    q, X, Y, Z = 9, 6, 1, 4
    bc = Boundary(type=[UNSET] * 5 + [NOSLIP], vel={DOWN: (0, 0, 0)})
    st = orc.lit_state(q, X, Y, Z)
    st["n"][...] = np.random.default_rng(1).random(st["n"].shape)
    b = st["n"].copy()
    orc.lit_noslip(st, bc)
    for x in range(2, X + 1):
        assert st["n"][1, 1, x, 1] == b[1, 1, x - 1, 2]
    for x in range(1, X):
        assert st["n"][1, 1, x, 2] == b[1, 1, x, 2]


def test_periodic_copies(orc):
    q, X, Y, Z = 19, 4, 3, 2
    c, _, _, _ = orc.tables(q)
    bc = Boundary(type=[PERIODIC] * 4 + [UNSET, UNSET])
    st = orc.lit_state(q, X, Y, Z)
    st["n"][...] = np.random.default_rng(2).random(st["n"].shape)
    b = st["n"].copy()
    orc.lit_periodic(st, bc)
    for k in range(1, q):
        if c[k][1] == 0 and c[k][0] > 0:
            assert np.array_equal(st["n"][1:-1, 1:-1, 1, k], b[1:-1, 1:-1, X + 1, k])
        if c[k][1] == 0 and c[k][0] < 0:
            assert np.array_equal(st["n"][1:-1, 1:-1, X, k], b[1:-1, 1:-1, 0, k])
        if c[k][1] > 0:  # y pass runs last and wins
            assert np.array_equal(st["n"][1:-1, 1, 1:-1, k], b[1:-1, Y + 1, 1:-1, k])
    assert np.array_equal(st["n"][0], b[0])  # no z periodicity (Boundary.cpp:91-99)


# ---------------------------------------------------------- stale quirks --
def _full_step(orc, st, bc, tau, F):
    orc.lit_collide(st, tau, F)
    orc.lit_push(st)
    assert orc.lit_noslip(st, bc) == 0
    orc.lit_periodic(st, bc)
    orc.lit_moments(st)
    orc.lit_reset(st, bc)


def test_stale_seam_next_to_wall(orc):
    """A fluid cell at a periodic seam next to a wall reads a never-pushed wall
    population: 0 after step 1, then (1-w) feq(rho0,u0) + w feq(rho_s, u_b + tau F/rho_s)."""
    q, X, Y, Z = 9, 6, 1, 5
    c, w, _, rev = orc.tables(q)
    bc = channel_bc(periodic_y=False)
    tau, F = 0.8, np.array([1e-3, 0.0, 0.0])
    st = orc.lit_state(q, X, Y, Z)
    orc.lit_init(st, 1.0, (0, 0, 0), None, bc)
    k_in = [k for k in range(q) if tuple(c[k]) == (-1, 0, 1)][0]  # arrives at (X, z=2)
    k_in2 = [k for k in range(q) if tuple(c[k]) == (1, 0, 1)][0]  # arrives at (1, z=2)
    _full_step(orc, st, bc, tau, F)
    assert st["n"][2, 1, X, k_in] == 0.0
    assert st["n"][2, 1, 1, k_in2] == 0.0
    _full_step(orc, st, bc, tau, F)
    om = 1.0 / tau
    kw = rev[k_in]
    exp = (1 - om) * w[kw] + om * feq_np(c.astype(float), w, 1.0, tau * F)[kw]
    assert st["n"][2, 1, X, k_in] == pytest.approx(exp, rel=1e-15)


def test_doubly_periodic_corner_never_wrapped(orc):
    """D3Q19, periodic x and y: the (+1,+1,0) population of cell (1,1) comes
    from an unwritten corner halo: 0 after step 1, feq(rho0,u0) after step 2."""
    q, X, Y, Z = 19, 4, 4, 5
    c, w, _, _ = orc.tables(q)
    bc = channel_bc()
    st = orc.lit_state(q, X, Y, Z)
    orc.lit_init(st, 1.0, (0, 0, 0), None, bc)
    k = [k for k in range(q) if tuple(c[k]) == (1, 1, 0)][0]
    _full_step(orc, st, bc, 0.7, (1e-4, 0, 0))
    assert np.all(st["n"][2:Z, 1, 1, k] == 0.0)
    _full_step(orc, st, bc, 0.7, (1e-4, 0, 0))
    assert np.all(st["n"][2:Z, 1, 1, k] == w[k])


def test_literal_corrected_differ_only_at_seams(orc):
    """After one step the literal and corrected populations differ only where
    a seam meets a wall or at doubly periodic corners."""
    q, X, Y, Z = 19, 6, 5, 6
    c, _, _, _ = orc.tables(q)
    bc = channel_bc()
    lit = orc.lit_state(q, X, Y, Z)
    orc.lit_init(lit, 1.0, (0, 0, 0), (1, 9, 1e-3), bc)
    orc.lit_moments(lit)
    orc.lit_reset(lit, bc)
    cor = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(cor, 1.0, (0, 0, 0), (1, 9, 1e-3))
    _full_step(orc, lit, bc, 0.7, (1e-5, 0, 0))
    orc.cor_run(cor, 0.7, (1e-5, 0, 0), bc, 1)
    diff = lit["n"][INTERIOR] != cor["f"]
    fl = orc.cor_flags(q, X, Y, Z, bc)
    diff[fl != 0] = False
    assert diff.any()
    for (z, y, x, k) in np.argwhere(diff):
        sx, sy = x - c[k][0], y - c[k][1]  # 0-based source cell
        seam = not (0 <= sx < X) or not (0 <= sy < Y)
        near_wall = z in (1, Z - 2) or (not (0 <= sx < X) and not (0 <= sy < Y))
        assert seam and near_wall, (z, y, x, k)


# ------------------------------------------------------------- moments ---
@pytest.mark.parametrize("q", [9, 19, 27])
def test_moments_of_feq(orc, q):
    X, Y, Z = 3, 1 if q == 9 else 2, 2
    st = orc.lit_state(q, X, Y, Z)
    rng = np.random.default_rng(q + 1)
    exp_rho = np.zeros(st["rho"].shape)
    exp_u = np.zeros(st["u"].shape)
    for idx in np.ndindex(st["rho"].shape):
        rho = rng.uniform(0.8, 1.2)
        u = rng.uniform(-0.1, 0.1, 3)
        if q == 9:
            u[1] = 0
        st["n"][idx] = orc.feq(q, rho, u)
        exp_rho[idx], exp_u[idx] = rho, u
    orc.lit_moments(st)
    assert np.allclose(st["rho"][INTERIOR], exp_rho[INTERIOR], rtol=1e-15)
    assert np.allclose(st["u"][INTERIOR], exp_u[INTERIOR], rtol=1e-13, atol=1e-16)


# --------------------------------------------------------- whole step -----
def test_closed_box_literal_equals_corrected(orc):
    """Lid-driven cavity (no seams): the staged reference step and the fused
    corrected step agree bit for bit on every fluid cell."""
    q, X, Y, Z = 19, 8, 7, 6
    types = [NOSLIP] * 6
    vel = {f: (0, 0, 0) for f in range(6)}
    vel[UP] = (0.05, 0.02, 0.0)
    bc = Boundary(type=types, vel=vel)
    lit = orc.lit_state(q, X, Y, Z)
    orc.lit_init(lit, 1.0, (0, 0, 0), (1, 4, 1e-3), bc)
    orc.lit_moments(lit)
    orc.lit_reset(lit, bc)
    cor = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(cor, 1.0, (0, 0, 0), (1, 4, 1e-3))
    orc.lit_run(lit, 0.75, (1e-5, 0, 0), bc, 30)
    orc.cor_run(cor, 0.75, (1e-5, 0, 0), bc, 30)
    rho, u = orc.cor_moments(cor, bc)
    fluid = orc.cor_flags(q, X, Y, Z, bc) == 0
    assert np.array_equal(lit["rho"][INTERIOR][fluid], rho[fluid])
    assert np.array_equal(lit["u"][INTERIOR][fluid], u[fluid])
    assert np.array_equal(lit["n"][INTERIOR][fluid], cor["f"][fluid])


def test_corrected_conserves_mass_literal_does_not(orc):
    q, X, Y, Z = 19, 8, 6, 8
    bc = channel_bc()
    cor = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(cor, 1.0, (0, 0, 0), (1, 3, 1e-3))
    fl = orc.cor_flags(q, X, Y, Z, bc) == 0
    m0 = orc.cor_populations(cor)[fl].sum()
    orc.cor_run(cor, 0.6, (1e-5, 0, 0), bc, 100)
    m1 = orc.cor_populations(cor)[fl].sum()
    assert abs(m1 - m0) <= 1e-12 * m0
    lit = orc.lit_state(q, X, Y, Z)
    orc.lit_init(lit, 1.0, (0, 0, 0), (1, 3, 1e-3), bc)
    orc.lit_moments(lit)
    orc.lit_reset(lit, bc)
    l0 = lit["n"][INTERIOR][fl].sum()
    orc.lit_run(lit, 0.6, (1e-5, 0, 0), bc, 100)
    l1 = lit["n"][INTERIOR][fl].sum()
    assert abs(l1 - l0) > 1e-8 * l0  # stale seam reads leak mass


def poiseuille_series(yt, H, g, nu, t, nmax=199):
    us = g / (2 * nu) * yt * (H - yt)
    s = np.zeros_like(yt)
    for n in range(1, nmax + 1, 2):
        s += 4 * g * H * H / (nu * math.pi ** 3 * n ** 3) * np.sin(n * math.pi * yt / H) * \
            math.exp(-n * n * math.pi ** 2 * nu * t / (H * H))
    return us - s


@pytest.mark.parametrize("mode", ["corrected", "literal"])
def test_poiseuille_transient_config1(orc, mode):
    """D2Q9 256x64, tau 0.8, g 1e-6, 10,000 steps: the x-averaged physical
    velocity u + F/(2 rho) follows the transient Fourier series with walls
    halfway between the wall layer and the first fluid row (H = 62)."""
    X, Y, Z, tau, g, steps = 256, 1, 64, 0.8, 1e-6, 10000
    bc = channel_bc(periodic_y=False)
    if mode == "corrected":
        st = orc.cor_state("f64", 9, X, Y, Z)
        orc.cor_init(st)
        orc.cor_run(st, tau, (g, 0, 0), bc, steps)
        rho, u = orc.cor_moments(st, bc)
        ux = (u[1:-1, 0, :, 0] + g / (2 * rho[1:-1, 0, :])).mean(axis=1)
    else:
        st = orc.lit_state(9, X, Y, Z)
        orc.lit_init(st, 1.0, (0, 0, 0), None, bc)
        orc.lit_run(st, tau, (g, 0, 0), bc, steps)
        rho, u = st["rho"], st["u"]
        cols = slice(1 + 16, X + 1 - 16)  # away from the seam columns
        ux = (u[2:Z, 1, cols, 0] + g / (2 * rho[2:Z, 1, cols])).mean(axis=1)
    H, nu = Z - 2, (tau - 0.5) / 3
    yt = np.arange(H) + 0.5
    ua = poiseuille_series(yt, H, g, nu, steps)
    err = np.max(np.abs(ux - ua)) / ua.max()
    assert err < 5e-3, err
    assert np.allclose(ux, ux[::-1], rtol=1e-9)  # symmetric channel


def test_couette_linear_profile(orc):
    """Moving lid, no force: steady profile is linear between the halfway walls."""
    q, X, Y, Z, tau, U = 9, 8, 1, 12, 1.1, 0.02
    bc = Boundary(type=[PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP],
                  vel={UP: (U, 0, 0), DOWN: (0, 0, 0)})
    st = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(st)
    orc.cor_run(st, tau, (0, 0, 0), bc, 8000)
    _, u = orc.cor_moments(st, bc)
    ux = u[1:-1, 0, :, 0].mean(axis=1)
    H = Z - 2
    exp = U * (np.arange(H) + 0.5) / H
    assert np.max(np.abs(ux - exp)) < 1e-3 * U


def test_taylor_green_decay(orc):
    """Fully periodic D3Q19: kinetic energy of the low-amplitude Taylor-Green
    field decays as exp(-2 nu |k|^2 t) with |k|^2 = 3 (2 pi / L)^2."""
    q, L, tau, steps = 19, 24, 0.8, 300
    bc = Boundary(type=[PERIODIC] * 6)
    st = orc.cor_state("f64", q, L, L, L)
    orc.cor_init(st, 1.0, (1e-3, 0, 0), (2, 0, 0.0))
    _, u0 = orc.cor_moments(st, bc)
    e0 = (u0 ** 2).sum()
    orc.cor_run(st, tau, (0, 0, 0), bc, steps)
    _, u1 = orc.cor_moments(st, bc)
    e1 = (u1 ** 2).sum()
    nu = (tau - 0.5) / 3
    k2 = 3 * (2 * math.pi / L) ** 2
    rate = -math.log(e1 / e0) / steps
    assert rate == pytest.approx(2 * nu * k2, rel=2e-2)


def test_axis_permutation_symmetry(orc):
    """Swapping x and y (and the matching populations) commutes with the step."""
    q, X, Y, Z = 19, 6, 5, 7
    c, _, _, _ = orc.tables(q)
    perm = [next(j for j in range(q) if tuple(c[j]) == (c[k][1], c[k][0], c[k][2])) for k in range(q)]
    bc = channel_bc()
    a = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(a, 1.0, (0, 0, 0), (1, 8, 1e-3))
    b = orc.cor_state("f64", q, Y, X, Z)
    b["f"][...] = np.ascontiguousarray(a["f"].transpose(0, 2, 1, 3)[..., np.argsort(perm)])
    orc.cor_run(a, 0.7, (1e-5, 0, 0), bc, 20)
    orc.cor_run(b, 0.7, (0, 1e-5, 0), bc, 20)
    bb = b["f"].transpose(0, 2, 1, 3)[..., perm]
    assert np.allclose(a["f"], bb, rtol=1e-13, atol=1e-16)
