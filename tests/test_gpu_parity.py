"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Literal mode (the reference's staged step) must match the oracle bit for bit
in FP64 and in the reference's FP32 rounding, after every stage.  Fused mode
must match the oracle's corrected step bit for bit for every storage variant
(same expression order on both sides, -ffp-contract=off on both compilers).
"""
import numpy as np
import pytest

import plbpd_amd as P
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, UNSET, UP, DOWN, NORTH, SOUTH, EAST, WEST

pytestmark = pytest.mark.gpu

LITERAL_CASES = [  # q, X, Y, Z
    (9, 8, 1, 6),
    (19, 5, 4, 3),
    (19, 8, 6, 5),
    (27, 5, 4, 3),
    (27, 7, 5, 6),
]


def both_bc(types, vel, rho_s=1.0):
    """The same boundary for the oracle and the library."""
    return Boundary(type=list(types), vel=dict(vel), rho_s=rho_s), P.boundary(types, vel, rho_s)


def couette_types(q):
    if q == 9:
        return [PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP]
    return [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP]


def lit_compare(st, orc_st, where):
    q = orc_st["q"]
    for name, field in (("rho", P.FIELD_RHO), ("u", P.FIELD_U), ("n", P.FIELD_N), ("ntmp", P.FIELD_NTMP)):
        dt = orc_st[name].dtype
        got = st.download(field, P.LAYOUT_REFERENCE, dtype=dt)
        exp = orc_st[name]
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            raise AssertionError(f"{where}: field {name} (q={q}) differs at {len(bad)} entries, first "
                                 f"{bad[:5].tolist()}: got {got[tuple(bad[0])]!r} expected {exp[tuple(bad[0])]!r}")


@pytest.mark.parametrize("q,X,Y,Z", LITERAL_CASES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_literal_stages_bitexact(orc, q, X, Y, Z, prec):
    """Every stage of PLBPD::march_in_time, 3 steps, moving lid, perturbed start."""
    print(f"\n[case] literal stages q={q} {X}x{Y}x{Z} {prec}", flush=True)
    dt = np.float64 if prec == "f64" else np.float32
    types = couette_types(q)
    ob, pb = both_bc(types, {UP: (0.05, 0.01 if q != 9 else 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)})
    tau, F = 0.7, (1e-4, -2e-5 if q != 9 else 0.0, 3e-5)
    o = orc.lit_state(q, X, Y, Z, dt)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 1234, 1e-2), ob)
    with P.State(q, X, Y, Z, prec, P.LITERAL) as st:
        st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 1234, 1e-2), pb)
        lit_compare(st, o, "initialize")
        for step in range(3):
            orc.lit_collide(o, tau, F)
            st.collide(tau, F)
            lit_compare(st, o, f"step {step} collide")
            orc.lit_push(o)
            st.stream("push")
            lit_compare(st, o, f"step {step} push")
            assert orc.lit_noslip(o, ob) == 0
            st.apply_noslip(pb)
            lit_compare(st, o, f"step {step} noslip")
            orc.lit_periodic(o, ob)
            st.apply_periodicity(pb)
            lit_compare(st, o, f"step {step} periodic")
            orc.lit_moments(o)
            st.calc_moments()
            lit_compare(st, o, f"step {step} moments")
            orc.lit_reset(o, ob)
            st.reset(pb)
            lit_compare(st, o, f"step {step} reset")


@pytest.mark.parametrize("q,X,Y,Z", LITERAL_CASES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_literal_collide_then_push_bitexact(orc, q, X, Y, Z, prec):
    """The staged calls with nothing between collide and push, so the library
    runs them as one kernel: n, ntmp, rho and u after every step equal the
    oracle's separate collide and push (ntmp holds the post-collision values
    the push read)."""
    print(f"\n[case] literal collide+push q={q} {X}x{Y}x{Z} {prec}", flush=True)
    dt = np.float64 if prec == "f64" else np.float32
    types = couette_types(q)
    ob, pb = both_bc(types, {UP: (0.05, 0.01 if q != 9 else 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)})
    tau, F = 0.7, (1e-4, -2e-5 if q != 9 else 0.0, 3e-5)
    o = orc.lit_state(q, X, Y, Z, dt)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 99, 1e-2), ob)
    with P.State(q, X, Y, Z, prec, P.LITERAL) as st:
        st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 99, 1e-2), pb)
        for step in range(3):
            orc.lit_collide(o, tau, F)
            orc.lit_push(o)
            st.collide(tau, F)
            st.stream("push")
            lit_compare(st, o, f"step {step} collide+push")
            assert orc.lit_noslip(o, ob) == 0
            orc.lit_periodic(o, ob)
            orc.lit_moments(o)
            orc.lit_reset(o, ob)
            st.apply_noslip(pb)
            st.apply_periodicity(pb)
            st.calc_moments()
            st.reset(pb)
            lit_compare(st, o, f"step {step} end")


@pytest.mark.parametrize("q,X,Y,Z", LITERAL_CASES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_literal_staged_steps_without_reads_bitexact(orc, q, X, Y, Z, prec):
    """Whole steps of staged calls with nothing read in between, the call
    sequence of PLBPD::march_in_time: the library folds each step's moments
    and resets into the next collide+push kernel.  The resets get a boundary
    of their own (another lid velocity and solid density than no-slip's), so
    the fold must take the reset calls' arguments."""
    print(f"\n[case] literal staged steps q={q} {X}x{Y}x{Z} {prec}", flush=True)
    dt = np.float64 if prec == "f64" else np.float32
    types = couette_types(q)
    ob, pb = both_bc(types, {UP: (0.05, 0.01 if q != 9 else 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)})
    obr, pbr = both_bc(types, {UP: (0.02, 0.0, 0.004), DOWN: (0.0, 0.003, 0.0)}, rho_s=1.01)
    tau, F = 0.7, (1e-4, -2e-5 if q != 9 else 0.0, 3e-5)
    o = orc.lit_state(q, X, Y, Z, dt)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 31, 1e-2), ob)
    with P.State(q, X, Y, Z, prec, P.LITERAL) as st:
        st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 31, 1e-2), pb)
        for step in range(4):
            orc.lit_collide(o, tau, F)
            orc.lit_push(o)
            assert orc.lit_noslip(o, ob) == 0
            orc.lit_periodic(o, ob)
            orc.lit_moments(o)
            orc.lit_reset(o, obr)
            st.collide(tau, F)
            st.stream("push")
            st.apply_noslip(pb)
            st.apply_periodicity(pb)
            st.calc_moments()
            st.reset(pbr)
        lit_compare(st, o, "after 4 staged steps")


@pytest.mark.parametrize("q,X,Y,Z", [(9, 1, 1, 1), (9, 3, 1, 2), (19, 1, 1, 1), (19, 2, 1, 3), (27, 1, 1, 1),
                                     (27, 1, 3, 2)])
@pytest.mark.parametrize("staged", [False, True])
def test_literal_tiny_domains_all_walls_bitexact(orc, q, X, Y, Z, staged):
    """Domains where every cell lies on several faces, all six faces no-slip
    with a velocity of their own: the face resets of each step (folded into
    the next step's kernel) must leave the last face's velocity, as the
    reference's face loop does.  Through plb_step and through the staged
    calls."""
    print(f"\n[case] literal tiny q={q} {X}x{Y}x{Z} staged={staged}", flush=True)
    types = [NOSLIP] * 6
    vel = {EAST: (0.0, 0.01, 0.0), WEST: (0.0, -0.01, 0.002), NORTH: (0.02, 0.0, 0.0),
           SOUTH: (-0.01, 0.0, 0.003), UP: (0.03, 0.0, 0.0), DOWN: (0.0, 0.004, 0.0)}
    ob, pb = both_bc(types, vel, rho_s=0.99)
    tau, F = 0.75, (1e-4, 0.0, -2e-5)
    o = orc.lit_state(q, X, Y, Z)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 3, 1e-2), ob)
    with P.State(q, X, Y, Z, "f64", P.LITERAL) as st:
        st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 3, 1e-2), pb)
        if staged:
            for _ in range(5):
                orc.lit_collide(o, tau, F)
                orc.lit_push(o)
                assert orc.lit_noslip(o, ob) == 0
                orc.lit_periodic(o, ob)
                orc.lit_moments(o)
                orc.lit_reset(o, ob)
                st.collide(tau, F)
                st.stream("push")
                st.apply_noslip(pb)
                st.apply_periodicity(pb)
                st.calc_moments()
                st.reset(pb)
        else:
            orc.lit_run(o, tau, F, ob, 5)
            st.configure(tau, F, pb)
            st.step(5)
        lit_compare(st, o, "after 5 steps")


def test_literal_deferred_collision_runs_before_other_calls(orc):
    """A literal collision waits only for a directly following push.  Two
    collisions in a row both run; a stream request the literal mode refuses
    ("pull") still finds the collision done, as in the reference, where
    the collision ran before Streaming threw (Streaming.cpp:96-102)."""
    q, X, Y, Z = 19, 6, 5, 4
    ob, pb = both_bc(couette_types(q), {UP: (0.03, 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)})
    tau, F = 0.8, (2e-5, 0.0, 0.0)
    o = orc.lit_state(q, X, Y, Z)
    orc.lit_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, 5, 1e-2), ob)
    with P.State(q, X, Y, Z, "f64", P.LITERAL) as st:
        st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 5, 1e-2), pb)
        orc.lit_collide(o, tau, F)
        orc.lit_collide(o, tau, F)
        orc.lit_push(o)
        st.collide(tau, F)
        st.collide(tau, F)
        st.stream("push")
        lit_compare(st, o, "collide, collide, push")
        orc.lit_collide(o, tau, F)
        st.collide(tau, F)
        with pytest.raises(P.LogicError):
            st.stream("pull")
        lit_compare(st, o, "collide, refused pull")
        orc.lit_push(o)
        st.stream("push")
        lit_compare(st, o, "then push")
        # the fusion knob turned off with a collision queued: it runs alone
        orc.lit_collide(o, tau, F)
        orc.lit_push(o)
        st.collide(tau, F)
        st.set_tuning(P.capi.TUNE_LITERAL_FUSE, 0)
        st.stream("push")
        lit_compare(st, o, "knob off with a queued collision")


@pytest.mark.parametrize("q,X,Y,Z", [(19, 9, 7, 6), (27, 7, 6, 8)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_literal_ntmp_after_folded_push(orc, q, X, Y, Z, prec):
    """The collide+push kernel stores in-place values on the outer layer
    only; ntmp elsewhere is rebuilt from n before anything reads it or
    changes n.  After folded steps: a collision that runs alone (a read
    follows it), a push alone, and an upload of n must each leave ntmp as the
    reference's separate stages do."""
    print(f"\n[case] literal ntmp after folded push q={q} {X}x{Y}x{Z} {prec}", flush=True)
    dt = np.float64 if prec == "f64" else np.float32
    ob, pb = both_bc(couette_types(q), {UP: (0.04, 0.01, 0.0), DOWN: (0.0, 0.0, 0.0)})
    tau, F = 0.66, (1e-4, -2e-5, 1e-5)
    o = orc.lit_state(q, X, Y, Z, dt)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 8, 1e-2), ob)
    with P.State(q, X, Y, Z, prec, P.LITERAL) as st:
        st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 8, 1e-2), pb)
        for _ in range(2):  # folded: collide, push, and the step before's moments and resets
            orc.lit_collide(o, tau, F)
            orc.lit_push(o)
            assert orc.lit_noslip(o, ob) == 0
            orc.lit_periodic(o, ob)
            orc.lit_moments(o)
            orc.lit_reset(o, ob)
            st.collide(tau, F)
            st.stream("push")
            st.apply_noslip(pb)
            st.apply_periodicity(pb)
            st.calc_moments()
            st.reset(pb)
        orc.lit_collide(o, tau, F)
        st.collide(tau, F)
        lit_compare(st, o, "collision run alone after a folded push")
        orc.lit_push(o)
        st.stream("push")
        lit_compare(st, o, "push alone")
        orc.lit_collide(o, tau, F)
        orc.lit_push(o)
        st.collide(tau, F)
        st.stream("push")
        new_n = (o["n"] * (1 + 1e-3 * np.random.default_rng(2).standard_normal(o["n"].shape))).astype(dt)
        o["n"][...] = new_n
        st.upload(P.FIELD_N, new_n)
        lit_compare(st, o, "n uploaded after a folded push")
        orc.lit_collide(o, tau, F)
        orc.lit_push(o)
        st.collide(tau, F)
        st.stream("push")
        lit_compare(st, o, "folded push after the upload")


@pytest.mark.parametrize("q,X,Y,Z", LITERAL_CASES)
@pytest.mark.parametrize("fuse", [1, 0])
def test_literal_run_bitexact(orc, q, X, Y, Z, fuse):
    """plb_step (the whole loop body) over 40 steps, FP64, with collision and
    push in one kernel (the default) and in two."""
    print(f"\n[case] literal run q={q} {X}x{Y}x{Z} fuse={fuse}", flush=True)
    types = couette_types(q)
    ob, pb = both_bc(types, {UP: (0.02, 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)})
    tau, F = 0.9, (1e-5, 0.0, 0.0)
    o = orc.lit_state(q, X, Y, Z)
    orc.lit_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, 7, 1e-3), ob)
    orc.lit_run(o, tau, F, ob, 40)
    with P.State(q, X, Y, Z, "f64", P.LITERAL) as st:
        st.set_tuning(P.capi.TUNE_LITERAL_FUSE, fuse)
        st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 7, 1e-3), pb)
        st.configure(tau, F, pb)
        st.step(40)
        lit_compare(st, o, "after 40 steps")


def test_eqlbdist_golden_on_gpu():
    """TestEqlbDist.cpp:11-35: D2Q9 feq at rho=1, u=(0.5,0.5,-0.5) equals the
    reference's nine float literals exactly, here produced by the GPU init."""
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    exp = np.array(g["eqlbdist_d2q9"]["expected"], np.float32)
    with P.State(9, 1, 1, 1, "f32", P.LITERAL) as st:
        st.initialize(1.0, (0.5, 0.5, -0.5))
        n = st.download(P.FIELD_N, P.LAYOUT_REFERENCE, dtype=np.float32)
    assert np.array_equal(n[1, 1, 1], exp)
    assert np.array_equal(n[0, 0, 0], exp)  # halo cells too (Domain.cpp:52-63)


def test_literal_poiseuille_c1_parity(orc):
    """D2Q9 256x64 channel, FP64, 10,000 steps: GPU literal == oracle literal."""
    X, Y, Z = 256, 1, 64
    print("\n[case] c1 D2Q9 256x1x64 10000 steps", flush=True)
    ob, pb = both_bc([PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP],
                     {UP: (0, 0, 0), DOWN: (0, 0, 0)})
    tau, F, steps = 0.8, (1e-6, 0.0, 0.0), 10000
    o = orc.lit_state(9, X, Y, Z)
    orc.lit_init(o, 1.0, (0, 0, 0), None, ob)
    orc.lit_run(o, tau, F, ob, steps)
    with P.State(9, X, Y, Z, "f64", P.LITERAL) as st:
        st.initialize(1.0, (0, 0, 0), None, pb)
        st.configure(tau, F, pb)
        st.step(steps)
        rho = st.download(P.FIELD_RHO)
        u = st.download(P.FIELD_U)
    interior = (slice(1, -1), slice(1, -1), slice(1, -1))
    r_rel = np.max(np.abs(rho[interior] - o["rho"][interior])) / np.max(np.abs(o["rho"][interior]))
    u_rel = np.max(np.abs(u[interior] - o["u"][interior])) / np.max(np.abs(o["u"][interior]))
    print(f"c1 max rel diff rho {r_rel:.3e} u {u_rel:.3e}")
    assert r_rel <= 1e-12 and u_rel <= 1e-12
    assert np.array_equal(rho, o["rho"]) and np.array_equal(u, o["u"])


@pytest.mark.parametrize("N,steps", [(24, 200)])
def test_literal_c2_reduced_parity(orc, N, steps):
    """D3Q19 channel (periodic x and y, walls up/down), seeded 1e-4
    perturbation, tau 0.6, F 1e-7: GPU literal == oracle literal."""
    print(f"\n[case] c2 reduced {N}^3 {steps} steps", flush=True)
    ob, pb = both_bc(couette_types(19), {UP: (0, 0, 0), DOWN: (0, 0, 0)})
    tau, F = 0.6, (1e-7, 0.0, 0.0)
    o = orc.lit_state(19, N, N, N)
    orc.lit_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, 2024, 1e-4), ob)
    orc.lit_run(o, tau, F, ob, steps)
    with P.State(19, N, N, N, "f64", P.LITERAL) as st:
        st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 2024, 1e-4), pb)
        st.configure(tau, F, pb)
        st.step(steps)
        assert np.array_equal(st.download(P.FIELD_RHO), o["rho"])
        assert np.array_equal(st.download(P.FIELD_U), o["u"])
        assert np.array_equal(st.download(P.FIELD_N), o["n"])


FUSED_CASES = [  # q, X, Y, Z, walls?, solids?
    (9, 16, 1, 12, True, False),
    (19, 16, 12, 10, True, False),
    (19, 64, 8, 9, True, True),
    (19, 12, 10, 8, False, False),
    (27, 16, 10, 9, True, False),
    (27, 10, 9, 8, True, True),
    # rows of 128 to 512 cells also run the row pull kernel (one cell per lane)
    (19, 256, 4, 6, True, True),
    (19, 128, 5, 7, False, False),
    (27, 128, 4, 5, True, True),
    (27, 384, 3, 4, True, False),
    (9, 256, 1, 8, True, False),
    # tiny and ragged: a 1 x 1 column that wraps onto itself, a walled 2 x 3 x 3 box
    (19, 1, 1, 4, False, False),
    (27, 2, 3, 3, True, False),
]
SCHEMES = {"push": 1, "pull": 2}


def fused_bc(q, walls, lid=(0.03, 0.0, 0.0)):
    if walls:
        t = couette_types(q)
        v = {UP: lid, DOWN: (0.0, 0.0, 0.0)}
    else:
        t = [PERIODIC] * 6 if q != 9 else [PERIODIC, PERIODIC, UNSET, UNSET, PERIODIC, PERIODIC]
        v = {}
    return both_bc(t, v)


@pytest.mark.parametrize("q,X,Y,Z,walls,solids", FUSED_CASES)
@pytest.mark.parametrize("variant", ["f64", "f32", "f64a32", "f32s", "f16s"])
@pytest.mark.parametrize("scheme", ["push", "pull"])
def test_fused_bitexact(orc, q, X, Y, Z, walls, solids, variant, scheme):
    print(f"\n[case] fused {scheme} {variant} q={q} {X}x{Y}x{Z} walls={walls} solids={solids}", flush=True)
    ob, pb = fused_bc(q, walls, (0.03, 0.0, 0.0))  # moving lid (both schemes)
    tau, F = 0.62, (2e-5, 1e-6 if q != 9 else 0.0, -1e-6)
    mask = None
    if solids:
        rng = np.random.default_rng(5)
        mask = (rng.random((Z, Y, X)) < 0.1).astype(np.uint8)
    init = (P.INIT_PERTURBED, 99, 2e-3)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0.01, 0, 0), init)
    orc.cor_run(o, tau, F, ob, 25, mask)
    with P.State(q, X, Y, Z, variant, P.FUSED) as st:
        st.set_tuning(P.capi.TUNE_SCHEME, SCHEMES[scheme])
        st.set_tuning(P.capi.TUNE_VECTOR, 1 if X in (128, 256, 384, 512) else 0)
        st.set_boundary(pb, mask)
        st.initialize(1.0, (0.01, 0, 0), init, pb)
        st.configure(tau, F, pb)
        st.step(25)
        raw = st.download_raw()
        rho, u = st.moments()
    orho, ou = orc.cor_moments(o, ob, mask)
    fl = orc.cor_flags(q, X, Y, Z, ob, mask)
    fluid = fl == 0
    if variant == "f16s":
        assert np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid])
    else:
        assert np.array_equal(raw[fluid], o["f"][fluid])
    assert np.array_equal(rho, orho) and np.array_equal(u, ou)


@pytest.mark.parametrize("q,X,Y,Z", [(19, 256, 6, 5), (27, 256, 4, 4), (27, 384, 3, 3)])
@pytest.mark.parametrize("variant", ["f64", "f32", "f64a32", "f32s", "f16s"])
def test_default_widths_around_solids(orc, q, X, Y, Z, variant):
    """With the automatic kernel choice around interior solids (row kernel
    with 4 cells per lane for D3Q19 4-byte storage, 2 for 8-byte storage and
    for D3Q27; the cell kernel, both candidates loaded, for 2-byte storage),
    the fused step equals the oracle bit for bit."""
    print(f"\n[case] default width q={q} {X}x{Y}x{Z} {variant}", flush=True)
    ob, pb = fused_bc(q, True, lid=(0.02, 0.0, 0.0))
    mask = (np.random.default_rng(8).random((Z, Y, X)) < 0.1).astype(np.uint8)
    init = (P.INIT_PERTURBED, 21, 1e-3)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0.01, 0, 0), init)
    orc.cor_run(o, 0.65, (1e-5, 0, 0), ob, 12, mask)
    with P.State(q, X, Y, Z, variant, P.FUSED) as st:
        st.set_boundary(pb, mask)
        st.initialize(1.0, (0.01, 0, 0), init, pb)
        st.configure(0.65, (1e-5, 0, 0), pb)
        st.step(12)
        raw = st.download_raw()
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    if variant == "f16s":
        assert np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid])
    else:
        assert np.array_equal(raw[fluid], o["f"][fluid])


@pytest.mark.parametrize("q,X,Y,Z,variant,vec,block", [
    (19, 256, 6, 5, "f64", -1, 128),   # automatic width: 2 cells per lane, two 128-cell segments
    (27, 256, 4, 4, "f64", -1, 128),   # D3Q27 around solids: 2 cells per lane
    (19, 512, 4, 3, "f32s", -1, 256),  # 4 cells per lane, two 256-cell segments
    (19, 256, 6, 5, "f32s", 1, 128),   # four 64-cell segments, two of them all solid
    (27, 384, 4, 3, "f32", 2, 128),    # three 128-cell segments
])
def test_row_pull_all_solid_segment(orc, q, X, Y, Z, variant, vec, block):
    """A solid block covers whole wavefront segments of some rows, so those
    wavefronts have no fluid cell.  Their neighbours in the row still pull
    the diagonal populations those wavefronts loaded from the fluid rows
    around the block: the row kernel must publish segment edges for every
    wavefront of an existing row, fluid or not."""
    print(f"\n[case] all-solid segment q={q} {X}x{Y}x{Z} {variant} vec={vec} block={block}", flush=True)
    ob, pb = fused_bc(q, True, lid=(0.02, 0.0, 0.0))
    mask = (np.random.default_rng(23).random((Z, Y, X)) < 0.1).astype(np.uint8)
    mask[:, 2:4, 0:block] = 1
    init = (P.INIT_PERTURBED, 41, 1e-3)
    tau, F = 0.66, (1e-5, 2e-6, 0)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0.01, 0, 0), init)
    orc.cor_run(o, tau, F, ob, 10, mask)
    with P.State(q, X, Y, Z, variant, P.FUSED) as st:
        st.set_tuning(P.capi.TUNE_SCHEME, SCHEMES["pull"])
        st.set_tuning(P.capi.TUNE_VECTOR, vec)
        st.set_boundary(pb, mask)
        st.initialize(1.0, (0.01, 0, 0), init, pb)
        st.configure(tau, F, pb)
        st.step(10)
        raw = st.download_raw()
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    assert np.array_equal(raw[fluid], o["f"][fluid])


@pytest.mark.parametrize("variant", ["f64", "f32", "f32s", "f16s"])
@pytest.mark.parametrize("X,vec", [(64, 1), (128, 1), (128, 2), (256, 1), (256, 2), (256, 4), (384, 1), (384, 2),
                                   (512, 1), (512, 2), (512, 4)])
@pytest.mark.parametrize("nt", [0, 1])
def test_closed_box_vector_pull(orc, variant, X, vec, nt):
    """All six faces walls (the lid moving): bounce-back on x-walls happens per
    element inside the row pull kernel, whose x shifts cross lanes and (for
    rows of several wavefronts) segments, every row shape the kernel is built
    for; with and without non-temporal loads and stores."""
    print(f"\n[case] row pull {variant} X={X} vec={vec} nt={nt}", flush=True)
    q, Y, Z = 19, 6, 5
    types = [NOSLIP] * 6
    vel = {f: (0.0, 0.0, 0.0) for f in range(6)}
    vel[UP] = (0.04, -0.01, 0.0)  # a moving lid: write-time wall terms in the row kernel
    ob, pb = both_bc(types, vel)
    rng = np.random.default_rng(2)
    mask = (rng.random((Z, Y, X)) < 0.05).astype(np.uint8)
    for solids in (None, mask):
        o = orc.cor_state(variant, q, X, Y, Z)
        orc.cor_init(o, 1.0, (0.0, 0, 0), (P.INIT_PERTURBED, 6, 1e-3))
        orc.cor_run(o, 0.7, (1e-5, 2e-6, 0), ob, 15, solids)
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            st.set_tuning(P.capi.TUNE_SCHEME, 2)
            st.set_tuning(P.capi.TUNE_VECTOR, vec)
            st.set_tuning(P.capi.TUNE_NT_LOAD, nt)
            st.set_tuning(P.capi.TUNE_NT_STORE, nt)
            st.set_boundary(pb, solids)
            st.initialize(1.0, (0.0, 0, 0), (P.INIT_PERTURBED, 6, 1e-3), pb)
            st.configure(0.7, (1e-5, 2e-6, 0), pb)
            st.step(15)
            raw = st.download_raw()
        fluid = orc.cor_flags(q, X, Y, Z, ob, solids) == 0
        assert np.array_equal(raw[fluid], o["f"][fluid])


@pytest.mark.parametrize("variant", ["f64", "f32s", "f16s"])
@pytest.mark.parametrize("solids", [False, True])
@pytest.mark.parametrize("lid", [0.0, 0.03])
def test_tuning_knobs_keep_results(orc, variant, solids, lid):
    """Every launch-shape knob (XCD mapping, tile order, block width,
    non-temporal loads and stores, row or cell kernel, push or pull) gives the
    oracle's populations bit for bit."""
    q, X, Y, Z = 19, 128, 16, 8  # 8 tile rows: the XCD y-stripe mapping applies
    ob, pb = fused_bc(q, True, lid=(lid, 0.0, 0.0))
    mask = None
    if solids:
        mask = (np.random.default_rng(4).random((Z, Y, X)) < 0.1).astype(np.uint8)
    init = (P.INIT_PERTURBED, 17, 1e-3)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0, 0, 0), init)
    orc.cor_run(o, 0.6, (1e-5, 0, 0), ob, 9, mask)
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    T = P.capi
    combos = [{}, {T.TUNE_XCD: 1}, {T.TUNE_XCD: 2}, {T.TUNE_XCD: 2, T.TUNE_VECTOR: 0}, {T.TUNE_TILE_ORDER: 4}, {T.TUNE_TILE_ORDER: 8, T.TUNE_XCD: 1},
              {T.TUNE_BLOCK_X: 64}, {T.TUNE_NT_LOAD: 1, T.TUNE_NT_STORE: 0}, {T.TUNE_VECTOR: 0},
              {T.TUNE_VECTOR: 1}, {T.TUNE_VECTOR: 2, T.TUNE_XCD: 1}, {T.TUNE_SCHEME: 1},
              {T.TUNE_SCHEME: 1, T.TUNE_NT_STORE: 0, T.TUNE_XCD: 1},
              {T.TUNE_INDEX64: 1, T.TUNE_VECTOR: 0}, {T.TUNE_INDEX64: 1, T.TUNE_VECTOR: 0, T.TUNE_SCHEME: 1}]
    for knobs in combos:
        print(f"\n[case] knobs {variant} solids={solids} lid={lid} {knobs}", flush=True)
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            for k, v in knobs.items():
                st.set_tuning(k, v)
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0, 0, 0), init, pb)
            st.configure(0.6, (1e-5, 0, 0), pb)
            st.step(9)
            raw = st.download_raw()
        if variant == "f16s":
            assert np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid]), knobs
        else:
            assert np.array_equal(raw[fluid], o["f"][fluid]), knobs


@pytest.mark.parametrize("variant", ["f64", "f32s"])
def test_boundary_change_mid_run(orc, variant):
    """Solids are added and the lid velocity changes between steps: the pull step (whose
    state owes the last step's streaming) gives the push step's result and
    the oracle's, the pending streaming using the old walls and solids."""
    q, X, Y, Z = 19, 32, 12, 10
    ob1, pb1 = fused_bc(q, True, lid=(0.02, 0.0, 0.0))
    ob2, pb2 = fused_bc(q, True, lid=(-0.01, 0.01, 0.0))
    rng = np.random.default_rng(12)
    m1 = (rng.random((Z, Y, X)) < 0.08).astype(np.uint8)
    # solids are only added (a cell that turns fluid would have no populations)
    m2 = m1 | (rng.random((Z, Y, X)) < 0.05).astype(np.uint8)
    init, tau, F = (P.INIT_PERTURBED, 31, 1e-3), 0.7, (1e-5, 0, 0)
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0, 0, 0), init)
    orc.cor_run(o, tau, F, ob1, 7, m1)
    orc.cor_run(o, tau, F, ob2, 6, m2)
    fluid = orc.cor_flags(q, X, Y, Z, ob2, m2) == 0
    for scheme in (2, 1):
        print(f"\n[case] boundary change {variant} scheme={scheme}", flush=True)
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            st.set_tuning(P.capi.TUNE_SCHEME, scheme)
            st.set_boundary(pb1, m1)
            st.initialize(1.0, (0, 0, 0), init, pb1)
            st.configure(tau, F, pb1)
            st.step(7)
            st.set_boundary(pb2, m2)
            st.configure(tau, F, pb2)
            st.step(6)
            raw = st.download_raw()
        assert np.array_equal(raw[fluid], o["f"][fluid]), scheme


@pytest.mark.parametrize("variant", ["f64", "f32s", "f16s"])
def test_pull_push_same_trajectory_with_midrun_reads(orc, variant):
    """Pull and push steps give identical populations, and reading the state
    mid-run (which converts post- to pre-collision storage) changes nothing.
    Wall and solid cells hold no populations in fused mode (each kernel leaves
    its own values there), so the comparison is over fluid cells."""
    q, X, Y, Z = 19, 32, 12, 10
    ob, pb = fused_bc(q, True, lid=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(8)
    mask = (rng.random((Z, Y, X)) < 0.1).astype(np.uint8)
    fl = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    with P.State(q, X, Y, Z, variant, P.FUSED) as a, P.State(q, X, Y, Z, variant, P.FUSED) as b:
        a.set_tuning(P.capi.TUNE_SCHEME, 2)
        b.set_tuning(P.capi.TUNE_SCHEME, 1)
        for st in (a, b):
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 21, 1e-3), pb)
            st.configure(0.58, (1e-5, 0, 0), pb)
        for chunk in (1, 7, 2, 10):
            a.step(chunk)
            b.step(chunk)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl])
            ra, ua = a.moments()
            rb, ub = b.moments()
            assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
        a.step_graph(8, 4)
        b.step(8)
        assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl])


@pytest.mark.parametrize("variant", ["f64", "f32s"])
@pytest.mark.parametrize("stream_type", ["push", "pull"])
def test_dropin_call_sequence_fused(orc, variant, stream_type):
    """The reference call sequence (collide, stream, noslip, periodicity,
    moments, reset) on a fused state equals plb_step, with either stream type
    (a fused state streams "pull", which the reference leaves unimplemented)."""
    print(f"\n[case] fused call sequence {variant} stream={stream_type}", flush=True)
    q, X, Y, Z = 19, 16, 8, 10
    ob, pb = fused_bc(q, True)
    tau, F = 0.7, (1e-5, 0, 0)
    with P.State(q, X, Y, Z, variant, P.FUSED) as a, P.State(q, X, Y, Z, variant, P.FUSED) as b:
        for st in (a, b):
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 3, 1e-3), pb)
        b.configure(tau, F, pb)
        for _ in range(6):
            a.collide(tau, F)
            a.stream(stream_type)
            a.apply_noslip(pb)
            a.apply_periodicity(pb)
            a.calc_moments()
            a.reset(pb)
        b.step(6)
        fl = orc.cor_flags(q, X, Y, Z, ob) == 0
        assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl])
        a.set_tuning(P.capi.TUNE_SCHEME, SCHEMES["push"])
        with pytest.raises(P.LogicError):  # "pull" contradicts the push scheme knob
            a.stream("pull")


def test_graph_replay_matches_eager(orc):
    """hipGraph replays (with eager work and syncs between) == plain launches
    (over fluid cells: wall cells hold no populations in fused mode)."""
    q, X, Y, Z = 19, 32, 16, 12
    ob, pb = fused_bc(q, True)
    fl = orc.cor_flags(q, X, Y, Z, ob) == 0
    with P.State(q, X, Y, Z, "f32s", P.FUSED) as a, P.State(q, X, Y, Z, "f32s", P.FUSED) as b:
        for st in (a, b):
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 11, 1e-3), pb)
            st.configure(0.8, (1e-5, 0, 0), pb)
        for rep in range(4):
            a.step_graph(10, 10)
            b.step(10)
            a.synchronize()
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), f"replay {rep}"
            a.global_moments()  # eager work between replays
        a.step_graph(13, 4)  # remainder path
        b.step(13)
        assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl])


@pytest.mark.parametrize("variant", ["f64", "f32s", "f16s"])
def test_global_moments_fp64(orc, variant):
    q, X, Y, Z = 19, 20, 12, 10
    ob, pb = fused_bc(q, True, lid=(0.0, 0.0, 0.0))
    o = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, 5, 1e-3))
    orc.cor_run(o, 0.7, (1e-5, 0, 0), ob, 10)
    f = orc.cor_populations(o)
    fl = orc.cor_flags(q, X, Y, Z, ob)
    c = orc.tables(q)[0]
    fluid = f[fl == 0]
    exp = np.array([fluid.sum(), (fluid * c[:, 0]).sum(), (fluid * c[:, 1]).sum(), (fluid * c[:, 2]).sum()])
    with P.State(q, X, Y, Z, variant, P.FUSED) as st:
        st.set_boundary(pb)
        st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 5, 1e-3), pb)
        st.configure(0.7, (1e-5, 0, 0), pb)
        st.step(10)
        got = st.global_moments()
    assert abs(got[0] - exp[0]) <= 1e-12 * abs(exp[0])
    assert np.allclose(got[1:], exp[1:], rtol=1e-9, atol=1e-13)


def test_error_codes():
    with pytest.raises(P.InvalidArgument):
        P.State(19, 0, 4, 4)
    with pytest.raises(P.LogicError):
        P.State(15, 4, 4, 4)
    with pytest.raises(P.InvalidArgument):
        P.State(19, 4, 4, 4, "f16s", P.LITERAL)
    with P.State(19, 4, 4, 4, "f64", P.LITERAL) as st:
        with pytest.raises(P.LogicError):
            st.stream("pull")
        with pytest.raises(P.LogicError):
            st.stream("sideways")
        bad = P.boundary([PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP], {UP: (0, 0, 0)})
        with pytest.raises(P.OutOfRange):
            st.apply_noslip(bad)
    with P.State(19, 4, 4, 4, "f32", P.FUSED) as st:
        with pytest.raises(P.LogicError):  # periodic on one side only
            st.set_boundary(P.boundary([PERIODIC, NOSLIP, PERIODIC, PERIODIC, PERIODIC, PERIODIC],
                                       {WEST: (0, 0, 0)}))


@pytest.mark.parametrize("fuse", [1, 0])
def test_timers_accumulate(fuse):
    """Per-stage HIP-event times of the literal loop.  A collision run with
    its push in one kernel counts under collide; stream then stays at zero."""
    _, pb = fused_bc(19, True)
    with P.State(19, 8, 8, 8, "f64", P.LITERAL) as st:
        st.set_tuning(P.capi.TUNE_LITERAL_FUSE, fuse)
        st.initialize(1.0, (0, 0, 0), None, pb)
        st.configure(0.8, (1e-5, 0, 0), pb)
        st.enable_timers(True)
        st.step(5)
        t = st.stage_times()
    for k in ("collide", "noslip", "periodicity", "moments"):
        assert t[k] > 0.0
    assert (t["stream"] > 0.0) == (fuse == 0)
