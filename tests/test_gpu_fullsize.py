"""The benchmark configurations at their full sizes.

c1 (D2Q9 256x64, 10,000 steps) and c2 (D3Q19 128^3, 1,000 steps) against the
CPU oracle, bit for bit; c3 (D3Q19 256^3 with 10% solids, 10,000 steps) and
c4 (D3Q27 384^3 Taylor-Green) through their error bounds against the FP64
run on the GPU, which the parity tests tie to the oracle; c4's Taylor-Green
start bit for bit against the oracle on 384-cell rows."""
import math

import numpy as np
import pytest

import plbpd_amd as P
from oracle.pyoracle import channel_bc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["literal", "fused"])
def test_c1_fullsize(orc, mode):
    X, Y, Z, tau, g, steps = 256, 1, 64, 0.8, 1e-6, 10000
    ob, pb = channel_bc(periodic_y=False), P.channel(periodic_y=False)
    print(f"\n[case] c1 {mode} full size", flush=True)
    if mode == "literal":
        o = orc.lit_state(9, X, Y, Z)
        orc.lit_init(o, 1.0, (0, 0, 0), None, ob)
        orc.lit_run(o, tau, (g, 0, 0), ob, steps)
        with P.State(9, X, Y, Z, "f64", P.LITERAL) as st:
            st.initialize(1.0, (0, 0, 0), None, pb)
            st.configure(tau, (g, 0, 0), pb)
            st.step(steps)
            assert np.array_equal(st.download(P.FIELD_RHO), o["rho"])
            assert np.array_equal(st.download(P.FIELD_U), o["u"])
            assert np.array_equal(st.download(P.FIELD_N), o["n"])
    else:
        c = orc.cor_state("f64", 9, X, Y, Z)
        orc.cor_init(c)
        orc.cor_run(c, tau, (g, 0, 0), ob, steps)
        orho, ou = orc.cor_moments(c, ob)
        with P.State(9, X, Y, Z, "f64", P.FUSED) as st:
            st.set_boundary(pb)
            st.initialize(1.0, (0, 0, 0), None, pb)
            st.configure(tau, (g, 0, 0), pb)
            st.step(steps)
            rho, u = st.moments()
        assert np.array_equal(rho, orho) and np.array_equal(u, ou)


@pytest.mark.parametrize("mode", ["literal", "fused"])
def test_c2_parity(orc, mode):
    N = 128
    steps, tau, g = 1000, 0.6, 1e-7
    init = (P.INIT_PERTURBED, 2024, 1e-4)
    ob, pb = channel_bc(), P.channel()
    print(f"\n[case] c2 {mode} {N}^3 x {steps}", flush=True)
    if mode == "literal":
        o = orc.lit_state(19, N, N, N)
        orc.lit_init(o, 1.0, (0, 0, 0), init, ob)
        orc.lit_run(o, tau, (g, 0, 0), ob, steps)
        with P.State(19, N, N, N, "f64", P.LITERAL) as st:
            st.initialize(1.0, (0, 0, 0), init, pb)
            st.configure(tau, (g, 0, 0), pb)
            st.step(steps)
            rho, u = st.download(P.FIELD_RHO), st.download(P.FIELD_U)
        it = (slice(1, -1),) * 3
        rel_u = np.max(np.abs(u[it] - o["u"][it])) / np.max(np.abs(o["u"][it]))
        rel_r = np.max(np.abs(rho[it] - o["rho"][it])) / np.max(np.abs(o["rho"][it]))
        print(f"c2 literal max rel diff: rho {rel_r:.3e} u {rel_u:.3e}")
        assert rel_u <= 1e-12 and rel_r <= 1e-12
        assert np.array_equal(rho, o["rho"]) and np.array_equal(u, o["u"])
    else:
        c = orc.cor_state("f64", 19, N, N, N)
        orc.cor_init(c, 1.0, (0, 0, 0), init)
        orc.cor_run(c, tau, (g, 0, 0), ob, steps)
        orho, ou = orc.cor_moments(c, ob)
        with P.State(19, N, N, N, "f64", P.FUSED) as st:
            st.set_boundary(pb)
            st.initialize(1.0, (0, 0, 0), init, pb)
            st.configure(tau, (g, 0, 0), pb)
            st.step(steps)
            rho, u = st.moments()
        assert np.array_equal(rho, orho) and np.array_equal(u, ou)


def test_c3_fullsize_error_bound():
    """D3Q19 256^3, walls up/down, 10% Bernoulli solids (the benchmark's
    seed), tau 0.55, g 5e-7, 10,000 steps: FP32-shifted storage keeps the
    velocity L2 error against the FP64 run at or below 1e-5, and the FP64
    mass sums of both runs drift by less than 1e-10 (stationary walls conserve
    mass up to rounding)."""
    import bench
    print("\n[case] c3 full size: f32s vs f64, 256^3, 10,000 steps", flush=True)
    mask = bench.solid_mask()
    err = bench.errors_vs_f64(P, ["f32s"], mask, 10000)
    e = err["f32s"]
    print(f"c3 f32s: u L2 {e['u_l2_rel']:.3e}, rho-1 L2 {e['rho_dev_l2_rel']:.3e}, "
          f"mass drift {e['mass_drift_fp64']:.2e}; f64 u max {err['f64_reference']['u_max']:.3e}", flush=True)
    assert e["u_l2_rel"] <= 1e-5
    assert abs(e["mass_drift_fp64"]) < 1e-10
    assert err["f64_reference"]["u_max"] > 1e-5  # the flow developed


def _tg_run(variant, X, Y, Z, steps, tau=0.52, u0=0.02):
    bc = P.channel()
    with P.State(27, X, Y, Z, variant, P.FUSED) as st:
        st.set_boundary(bc)
        st.initialize(1.0, (u0, 0.0, 0.0), (P.INIT_TAYLOR_GREEN, 0, 0.0), bc)
        st.configure(tau, (0.0, 0.0, 0.0), bc)
        g0 = st.global_moments()
        rho0, u0f = st.moments()
        st.step(steps)
        g1 = st.global_moments()
        rho, u = st.moments()
    return {"g0": g0, "g1": g1, "rho0": rho0, "u0": u0f, "rho": rho, "u": u}


@pytest.mark.parametrize("variant", ["f64", "f32", "f32s", "f16s"])
def test_c4_taylor_green_start_bitexact(orc, variant):
    """c4's initial state (PLB_INIT_TAYLOR_GREEN: u = u0 (sin kx cos kz cos ky,
    0, -cos kx sin kz cos ky), k = 2 pi / L per axis, walls up/down) and 30
    corrected steps at tau 0.52 on D3Q27 rows of 384 cells, bit for bit against
    the oracle's start of the same kind and its steps."""
    X, Y, Z, steps = 384, 4, 6, 30
    print(f"\n[case] c4 Taylor-Green start {variant} {X}x{Y}x{Z}", flush=True)
    ob = channel_bc()
    o = orc.cor_state(variant, 27, X, Y, Z)
    orc.cor_init(o, 1.0, (0.02, 0.0, 0.0), (P.INIT_TAYLOR_GREEN, 0, 0.0))
    f_start = o["f"].copy()
    orc.cor_run(o, 0.52, (0.0, 0.0, 0.0), ob, steps)
    fluid = orc.cor_flags(27, X, Y, Z, ob) == 0
    bc = P.channel()
    with P.State(27, X, Y, Z, variant, P.FUSED) as st:
        st.set_boundary(bc)
        st.initialize(1.0, (0.02, 0.0, 0.0), (P.INIT_TAYLOR_GREEN, 0, 0.0), bc)
        start = st.download_raw()
        st.configure(0.52, (0.0, 0.0, 0.0), bc)
        st.step(steps)
        raw = st.download_raw()
        rho, u = st.moments()
    view = (lambda a: a.view(np.uint16)) if variant == "f16s" else (lambda a: a)
    assert np.array_equal(view(start), view(f_start))
    assert np.array_equal(view(raw)[fluid], view(o["f"])[fluid])
    orho, ou = orc.cor_moments(o, ob)
    assert np.array_equal(rho, orho) and np.array_equal(u, ou)
    assert np.abs(u[fluid]).max() > 1e-3  # the vortex is there


def test_c4_fullsize_kinetic_energy_and_mass():
    """D3Q27 384^3 Taylor-Green (u0 0.02, k = 2 pi / 384), walls up/down,
    tau 0.52, 1,000 steps in FP64, FP32-shifted, FP16-shifted and FP32
    storage.  Bounds: the kinetic energy of every reduced variant within
    1e-3 of the FP64 run's; FP64 mass sums drift by at most 1e-11 (FP64),
    1e-8 (FP32-shifted), 1e-6 (FP16-shifted), 1e-3 (FP32, which rounds the
    O(1) populations); the FP64 energy decays at least as fast as the
    wall-free periodic vortex, exp(-2 nu 3 k^2 t), and no more than 5% in
    1,000 steps."""
    L, steps = 384, 1000
    fluid = np.ones((L, L, L), bool)
    fluid[0] = fluid[-1] = False  # the wall layers
    drift_bound = {"f64": 1e-11, "f32s": 1e-8, "f16s": 1e-6, "f32": 1e-3}
    ke = {}
    for v in ("f64", "f32s", "f16s", "f32"):
        print(f"\n[case] c4 full size {v} {L}^3 x {steps}", flush=True)
        r = _tg_run(v, L, L, L, steps)
        ke0 = float(0.5 * (r["rho0"][fluid] * (r["u0"][fluid] ** 2).sum(-1)).sum())
        ke1 = float(0.5 * (r["rho"][fluid] * (r["u"][fluid] ** 2).sum(-1)).sum())
        drift = (r["g1"][0] - r["g0"][0]) / r["g0"][0]
        ke[v] = (ke0, ke1)
        print(f"c4 {v}: KE ratio {ke1 / ke0:.6f}, mass drift {drift:.3e}", flush=True)
        assert abs(drift) <= drift_bound[v], (v, drift)
        del r
    nu = (0.52 - 0.5) / 3
    periodic = math.exp(-2 * nu * 3 * (2 * math.pi / L) ** 2 * steps)
    ratio64 = ke["f64"][1] / ke["f64"][0]
    assert 0.95 < ratio64 <= periodic
    for v in ("f32s", "f16s", "f32"):
        assert abs(ke[v][1] / ke["f64"][1] - 1) <= 1e-3, v
        assert abs(ke[v][0] / ke["f64"][0] - 1) <= 1e-3, v


@pytest.mark.parametrize("X", [64, 128])  # cell kernel; row kernel (2 cells per lane)
def test_past_4gib_index_range(X):
    """A domain whose population buffers pass 2^32 bytes takes the 64-bit
    index kernels.  Interior solids repeating every 12 z-planes (12 does not
    divide the 2^17 planes of a 4 GiB address wrap), a uniform start and a
    body force give a flow with the same period, so after 20 steps every
    12-plane slab must equal, bit for bit, a 12-plane domain that takes the
    32-bit kernels."""
    Y, period, steps = 64, 12, 20
    Z = period * ((28311552 // (X * Y)) // period)
    q, esize = 19, 8
    assert q * X * Y * Z * esize > 2**32  # past the 32-bit byte-offset range
    rng = np.random.default_rng(12)
    tile = (rng.random((period, Y, X)) < 0.1).astype(np.uint8)
    # walls on the y faces; x and z periodic
    pb = P.boundary([P.PERIODIC, P.PERIODIC, P.NOSLIP, P.NOSLIP, P.PERIODIC, P.PERIODIC],
                    {P.NORTH: (0.0, 0.0, 0.0), P.SOUTH: (0.0, 0.0, 0.0)})
    print(f"\n[case] past 4 GiB: D3Q19 f64 {X}x{Y}x{Z}, solids period {period}", flush=True)

    def run(nz, mask):
        with P.State(q, X, Y, nz, "f64", P.FUSED) as st:
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0.01, 0, 0), None, pb)
            st.configure(0.6, (2e-6, 0, 0), pb)
            st.step(steps)
            return st.moments()

    rho_s, u_s = run(period, tile)
    rho_b, u_b = run(Z, np.tile(tile, (Z // period, 1, 1)))
    reps = Z // period
    assert np.array_equal(rho_b.reshape(reps, -1), np.broadcast_to(rho_s.reshape(1, -1), (reps, rho_s.size)))
    assert np.array_equal(u_b.reshape(reps, -1), np.broadcast_to(u_s.reshape(1, -1), (reps, u_s.size)))
    assert np.abs(u_s).max() > 1e-6  # the flow really moved
