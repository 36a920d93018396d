"""Parity with the CPU oracle at the BASELINE configurations' full sizes.

c1 (D2Q9 256x64, 10,000 steps) runs by default.  c2 at 128^3 x 1,000 steps
costs the oracle about two minutes on 16 host cores, so it runs only with
PLBPD_FULLSIZE=1; by default the same test runs at 64^3 x 1,000 steps."""
import os

import numpy as np
import pytest

import plbpd_amd as P
from oracle.pyoracle import channel_bc

pytestmark = pytest.mark.gpu
FULL = os.environ.get("PLBPD_FULLSIZE") == "1"


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
    N = 128 if FULL else 64
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
