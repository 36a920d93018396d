"""One-array states (PLB_MODE_FUSED_AA, the AA pattern): the fused step's
numbers with one population array, against the oracle's corrected step and
against the two-array pull step, bit for bit."""
import numpy as np
import pytest

import plbpd_amd as P
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, UNSET, UP, DOWN

pytestmark = pytest.mark.gpu

CASES = [  # q, X, Y, Z, walls, solids
    (9, 16, 1, 12, True, False),
    (19, 16, 12, 10, True, False),
    (19, 64, 8, 9, True, True),
    (19, 12, 10, 8, False, False),
    (27, 10, 9, 8, True, True),
    (19, 256, 4, 6, True, True),
    (19, 1, 1, 4, False, False),
    (27, 2, 3, 3, True, False),
]


def bcs(q, walls, lid=(0.03, 0.0, 0.0)):
    if walls:
        t = [PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP] if q == 9 else \
            [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP]
        v = {UP: lid, DOWN: (0.0, 0.0, 0.0)}
    else:
        t = [PERIODIC] * 6 if q != 9 else [PERIODIC, PERIODIC, UNSET, UNSET, PERIODIC, PERIODIC]
        v = {}
    return Boundary(type=t, vel=v), P.boundary(t, v)


def fluid_of(orc, ob, q, X, Y, Z, mask=None):
    """Cells whose populations the fused step defines: wall-face cells and
    solids hold no populations (the two-array step leaves stale values where a
    whole wavefront is not fluid; the one-array step stores zeros)."""
    return orc.cor_flags(q, X, Y, Z, ob, mask) == 0


@pytest.mark.parametrize("q,X,Y,Z,walls,solids", CASES)
@pytest.mark.parametrize("variant", ["f64", "f32", "f64a32", "f32s", "f16s"])
@pytest.mark.parametrize("steps", [24, 25])
def test_aa_matches_oracle(orc, q, X, Y, Z, walls, solids, variant, steps):
    """Even and odd step counts (an odd count is read through the gather into
    a temporary second array), moving lid, interior solids."""
    print(f"\n[case] aa {variant} q={q} {X}x{Y}x{Z} walls={walls} solids={solids} steps={steps}", flush=True)
    ob, pb = bcs(q, walls)
    tau, F = 0.62, (2e-5, 1e-6 if q != 9 else 0.0, -1e-6)
    mask = (np.random.default_rng(5).random((Z, Y, X)) < 0.1).astype(np.uint8) if solids else None
    init = (P.INIT_PERTURBED, 99, 2e-3)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0.01, 0, 0), init)
    orc.cor_run(o, tau, F, ob, steps, mask)
    with P.State(q, X, Y, Z, variant, P.FUSED_AA) as st:
        assert st.scheme() == "aa"
        st.set_boundary(pb, mask)
        st.initialize(1.0, (0.01, 0, 0), init, pb)
        st.configure(tau, F, pb)
        st.step(steps)
        raw = st.download_raw()
        rho, u = st.moments()
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    if variant == "f16s":
        assert np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid])
    else:
        assert np.array_equal(raw[fluid], o["f"][fluid])
    orho, ou = orc.cor_moments(o, ob, mask)
    assert np.array_equal(rho, orho) and np.array_equal(u, ou)


def test_aa_holds_one_population_array():
    with P.State(19, 32, 32, 32, "f32s", P.FUSED_AA) as a, P.State(19, 32, 32, 32, "f32s", P.FUSED) as b:
        one = b.nbytes() - a.nbytes()
        # the two-array state holds exactly one more population array
        assert one > 0 and one % (19 * 4) == 0 and one >= 19 * 4 * 32 ** 3


@pytest.mark.parametrize("variant", ["f64", "f32s", "f16s"])
def test_aa_equals_pull_with_midrun_reads_and_graphs(orc, variant):
    """Reads after odd step counts (gathered into a temporary array, then the
    state continues from one array), graph replays and remainders: the
    one-array state keeps the pull step's populations and moments."""
    q, X, Y, Z = 19, 32, 12, 10
    ob, pb = bcs(q, True, lid=(0.02, 0.0, 0.0))
    mask = (np.random.default_rng(8).random((Z, Y, X)) < 0.1).astype(np.uint8)
    fl = fluid_of(orc, ob, q, X, Y, Z, mask)
    with P.State(q, X, Y, Z, variant, P.FUSED_AA) as a, P.State(q, X, Y, Z, variant, P.FUSED) as b:
        for st in (a, b):
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 21, 1e-3), pb)
            st.configure(0.58, (1e-5, 0, 0), pb)
        for chunk in (1, 7, 2, 10, 3):
            a.step(chunk)
            b.step(chunk)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), chunk
            ra, ua = a.moments()
            rb, ub = b.moments()
            assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
            assert np.array_equal(a.global_moments(), b.global_moments())
        for rep in range(3):
            a.step_graph(9, 4)  # an odd state first runs one plain step; remainder 1 or 0
            b.step(9)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), rep


@pytest.mark.parametrize("stream_type", ["push", "pull"])
def test_aa_dropin_call_sequence(orc, stream_type):
    """The reference's staged calls on a one-array state, including a stream
    without a collision before it (the stream-only step), equal the pull
    state's."""
    q, X, Y, Z = 19, 16, 8, 10
    ob, pb = bcs(q, True)
    fl = fluid_of(orc, ob, q, X, Y, Z)
    tau, F = 0.7, (1e-5, 0, 0)
    with P.State(q, X, Y, Z, "f32s", P.FUSED_AA) as a, P.State(q, X, Y, Z, "f32s", P.FUSED) as b:
        for st in (a, b):
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 3, 1e-3), pb)
        for i in range(7):
            for st in (a, b):
                if i != 3:
                    st.collide(tau, F)
                st.stream(stream_type)
                st.apply_noslip(pb)
                st.apply_periodicity(pb)
                st.calc_moments()
                st.reset(pb)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), i


@pytest.mark.parametrize("steps_before", [6, 7])
def test_aa_boundary_change_mid_run(orc, steps_before):
    """Solids added and the lid changed after an even or an odd number of
    steps: the one-array state finishes the pending streaming with the old
    walls, as the pull state does."""
    q, X, Y, Z = 19, 32, 12, 10
    _, pb1 = bcs(q, True, lid=(0.02, 0.0, 0.0))
    ob2, pb2 = bcs(q, True, lid=(-0.01, 0.01, 0.0))
    rng = np.random.default_rng(12)
    m1 = (rng.random((Z, Y, X)) < 0.08).astype(np.uint8)
    m2 = m1 | (rng.random((Z, Y, X)) < 0.05).astype(np.uint8)
    with P.State(q, X, Y, Z, "f64", P.FUSED_AA) as a, P.State(q, X, Y, Z, "f64", P.FUSED) as b:
        for st in (a, b):
            st.set_boundary(pb1, m1)
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 31, 1e-3), pb1)
            st.configure(0.7, (1e-5, 0, 0), pb1)
            st.step(steps_before)
            st.set_boundary(pb2, m2)
            st.configure(0.7, (1e-5, 0, 0), pb2)
            st.step(6)
        fl = fluid_of(orc, ob2, q, X, Y, Z, m2)
        assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl])
        ra, ua = a.moments()
        rb, ub = b.moments()
        assert np.array_equal(ra, rb) and np.array_equal(ua, ub)


def test_aa_scheme_knob():
    with P.State(19, 8, 8, 8, "f32s", P.FUSED_AA) as a, P.State(19, 8, 8, 8, "f32s", P.FUSED) as b:
        with pytest.raises(P.LogicError):
            a.set_tuning(P.capi.TUNE_SCHEME, P.capi.SCHEME_PULL)
        with pytest.raises(P.LogicError):
            b.set_tuning(P.capi.TUNE_SCHEME, P.capi.SCHEME_AA)
        a.set_tuning(P.capi.TUNE_SCHEME, P.capi.SCHEME_AA)
        assert a.scheme() == "aa" and b.scheme() == "pull"
