"""Two pull steps per launch (PLB_TUNE_TWO_STEP): the numbers of the oracle's
corrected step, and of the one-step pull kernels, bit for bit, for every
storage variant the two-step kernel is built for; tiles cut ragged in y,
several z-chunks, periodic seams on every axis, walls on every axis (with a
moving lid), interior solids; even and odd step counts, reads between runs
and graph replays; and at the full 256^3 size of the headline workload."""
import numpy as np
import pytest

import plbpd_amd as P
from plbpd_amd import capi
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, EAST, WEST, NORTH, SOUTH, UP, DOWN

pytestmark = pytest.mark.gpu

VARIANTS = {19: ["f64", "f32", "f64a32", "f32s", "f16s"], 27: ["f64", "f32", "f32s", "f16s"]}

# q, X, Y, Z, walls ("z": up/down, "y": north/south, "x": east/west, None), solids, z-chunk
CASES = [
    (19, 64, 12, 10, "z", True, 3),
    (19, 64, 7, 9, "z", False, 0),
    (19, 128, 5, 6, None, True, 4),
    (19, 64, 10, 8, "y", True, 5),
    (19, 64, 9, 7, "x", False, 2),
    (19, 64, 1, 5, None, False, 1),
    (27, 64, 11, 9, "z", True, 4),
    (27, 64, 6, 6, "y", False, 0),
    (19, 96, 9, 7, "z", True, 3),
    # one and two planes (every z-neighbour is the cell's own plane or the
    # other one), and one fluid plane between z-walls
    (19, 64, 6, 1, None, True, 0),
    (19, 64, 5, 2, "y", True, 1),
    (27, 64, 4, 3, "z", True, 0),
]


def bcs(walls, lid=(0.03, 0.0, 0.0)):
    t = [PERIODIC] * 6
    v = {}
    if walls == "z":
        t[UP] = t[DOWN] = NOSLIP
        v = {UP: lid, DOWN: (0.0, 0.0, 0.0)}
    elif walls == "y":
        t[NORTH] = t[SOUTH] = NOSLIP
        v = {NORTH: (0.02, 0.0, -0.01), SOUTH: (0.0, 0.0, 0.0)}
    elif walls == "x":
        t[EAST] = t[WEST] = NOSLIP
        v = {EAST: (0.0, 0.02, 0.01), WEST: (0.0, 0.0, 0.0)}
    return Boundary(type=t, vel=v), P.boundary(t, v)


def two_step(st, on=True, zc=0, bx=0):
    st.set_tuning(capi.TUNE_TWO_STEP, 1 if on else 0)
    st.set_tuning(capi.TUNE_TWO_ZCHUNK, zc)
    st.set_tuning(capi.TUNE_TWO_TILE_X, bx)


@pytest.mark.parametrize("q,X,Y,Z,walls,solids,zc", CASES)
@pytest.mark.parametrize("steps", [12, 13])
def test_two_step_matches_oracle(orc, q, X, Y, Z, walls, solids, zc, steps):
    ob, pb = bcs(walls)
    tau, F = 0.61, (2e-5, -1e-6, 1e-6)
    mask = (np.random.default_rng(17).random((Z, Y, X)) < 0.1).astype(np.uint8) if solids else None
    init = (P.INIT_PERTURBED, 5, 2e-3)
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    for variant in VARIANTS[q]:
        print(f"\n[case] two-step {variant} q={q} {X}x{Y}x{Z} walls={walls} solids={solids} zc={zc} "
              f"steps={steps}", flush=True)
        o = orc.cor_state(variant, q, X, Y, Z)
        orc.cor_init(o, 1.0, (0.01, 0.0, 0.0), init)
        orc.cor_run(o, tau, F, ob, steps, mask)
        orho, ou = orc.cor_moments(o, ob, mask)
        # 64- and 32-cell tiles; per-population base pointers (the form for
        # arrays beyond 4 GiB, forced by the 64-bit index setting); workgroup
        # i on tile i instead of the XCD runs
        for bx, wide, xcd in ((64, 0, -1), (32, 0, -1), (64, 1, -1), (32, 1, -1), (64, 0, 0)):
            if X % bx:
                continue
            if q == 27 and variant == "f64" and bx == 64:
                continue  # its ring fits 32-cell tiles only
            with P.State(q, X, Y, Z, variant, P.FUSED) as st:
                two_step(st, True, zc, bx)
                st.set_tuning(capi.TUNE_INDEX64, wide)
                st.set_tuning(capi.TUNE_XCD, xcd)
                assert st.scheme() == "pull2", (variant, bx, wide)
                st.set_boundary(pb, mask)
                st.initialize(1.0, (0.01, 0.0, 0.0), init, pb)
                st.configure(tau, F, pb)
                st.step(steps)
                raw = st.download_raw()
                rho, u = st.moments()
            if variant == "f16s":
                assert np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid]), (variant, bx, wide, xcd)
            else:
                assert np.array_equal(raw[fluid], o["f"][fluid]), (variant, bx, wide, xcd)
            assert np.array_equal(rho, orho) and np.array_equal(u, ou), (variant, bx, wide, xcd)


@pytest.mark.parametrize("variant,bx", [("f64", 64), ("f32s", 64), ("f16s", 64), ("f32s", 32)])
def test_two_step_equals_one_step_with_reads_and_graphs(orc, variant, bx):
    """Runs split by reads (each read gathers the pending streaming), graph
    replays of 10 steps (four two-step launches and two one-step launches, so
    that the graph keeps its buffer parity) and of 12 (six two-step launches),
    and odd remainders."""
    q, X, Y, Z = 19, 64, 13, 11
    ob, pb = bcs("z", lid=(0.02, 0.0, 0.0))
    mask = (np.random.default_rng(4).random((Z, Y, X)) < 0.08).astype(np.uint8)
    fl = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    with P.State(q, X, Y, Z, variant, P.FUSED) as a, P.State(q, X, Y, Z, variant, P.FUSED) as b:
        two_step(a, True, 4, bx)
        two_step(b, False)
        for st in (a, b):
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 9, 1e-3), pb)
            st.configure(0.57, (1e-5, 0, 0), pb)
        for chunk in (1, 2, 7, 4, 3):
            a.step(chunk)
            b.step(chunk)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), chunk
            ra, ua = a.moments()
            rb, ub = b.moments()
            assert np.array_equal(ra, rb) and np.array_equal(ua, ub), chunk
        for spg in (10, 12, 10):
            a.step_graph(2 * spg + 1, spg)
            b.step(2 * spg + 1)
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), spg
        assert np.array_equal(a.global_moments(), b.global_moments())


def test_two_step_knobs():
    with P.State(19, 64, 8, 8, "f32s", P.FUSED) as st:
        with pytest.raises(P.InvalidArgument):
            st.set_tuning(capi.TUNE_TWO_STEP, 2)
        with pytest.raises(P.InvalidArgument):
            st.set_tuning(capi.TUNE_TWO_ZCHUNK, -1)
        with pytest.raises(P.InvalidArgument):
            st.set_tuning(capi.TUNE_TWO_TILE_X, 48)
        st.set_tuning(capi.TUNE_TWO_STEP, -1)
        st.set_tuning(capi.TUNE_TWO_ZCHUNK, 0)
        # automatic: not below 2^18 cells
        assert st.scheme() == "pull" and st.steps_per_launch() == 1
        st.set_tuning(capi.TUNE_TWO_STEP, 1)
        assert st.scheme() == "pull2" and st.steps_per_launch() == 2
        st.set_tuning(capi.TUNE_TWO_STEP, 0)
        assert st.scheme() == "pull"
    # automatic: 4- and 8-byte storage from 2^18 cells (2^19 for D3Q27);
    # forced on where a kernel exists
    with P.State(19, 64, 64, 64, "f16s", P.FUSED) as st:
        assert st.scheme() == "pull"
        st.set_tuning(capi.TUNE_TWO_STEP, 1)
        assert st.scheme() == "pull2"
    for q, dims, v, want in ((19, (64, 64, 64), "f64", "pull2"), (19, (64, 64, 32), "f32s", "pull"),
                             (27, (64, 64, 64), "f32s", "pull"), (27, (64, 64, 128), "f32s", "pull2")):
        with P.State(q, *dims, v, P.FUSED) as st:
            assert st.scheme() == want, (q, dims, v)
    for q, x, v in ((19, 80, "f32s"), (9, 64, "f32s")):
        with P.State(q, x, 1 if q == 9 else 8, 8, v, P.FUSED) as st:
            st.set_tuning(capi.TUNE_TWO_STEP, 1)
            assert st.scheme() == "pull", (q, x, v)
    # D3Q27 with 8-byte storage: 32-cell tiles only
    with P.State(27, 64, 8, 8, "f64", P.FUSED) as st:
        st.set_tuning(capi.TUNE_TWO_STEP, 1)
        assert st.scheme() == "pull2"
        st.set_tuning(capi.TUNE_TWO_TILE_X, 64)
        assert st.scheme() == "pull"
    with P.State(19, 64, 8, 8, "f32s", P.FUSED) as st:
        st.set_tuning(capi.TUNE_SCHEME, capi.SCHEME_PUSH)
        assert st.scheme() == "push"


@pytest.mark.parametrize("variant,solids,bx", [("f32s", True, 64), ("f32s", False, 64), ("f16s", True, 64),
                                               ("f64", True, 64), ("f32s", True, 32)])
def test_two_step_full_size(variant, solids, bx):
    """The headline lattice (D3Q19 256^3 channel, 10% solid voxels): two-step
    launches equal one-step launches bit for bit after 20 steps (graph replays,
    the launch geometry of the benchmark: eight XCD stripes, two z-chunks)."""
    N = 256
    bc = P.channel()
    mask = (np.random.default_rng(20250303).random((N, N, N)) < 0.10).astype(np.uint8) if solids else None
    print(f"\n[case] two-step full size {variant} solids={solids}", flush=True)
    out = []
    for on in (True, False):
        with P.State(19, N, N, N, variant, P.FUSED) as st:
            two_step(st, on, 0, bx)
            st.set_boundary(bc, mask)
            st.initialize(1.0, (0.0, 0.0, 0.0), (P.INIT_PERTURBED, 3, 1e-3), bc)
            st.configure(0.55, (5e-7, 0.0, 0.0), bc)
            st.step_graph(21, 20)
            out.append((st.download_raw(), st.global_moments()))
    fluid = np.ones((N, N, N), bool) if mask is None else mask == 0
    fluid[0] = False
    fluid[-1] = False
    assert np.array_equal(out[0][0][fluid], out[1][0][fluid])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("variant", ["f32s", "f64"])
def test_staged_calls_pair_into_two_step_launches(orc, variant):
    """The reference's staged calls (collide, stream, no-slip, periodicity,
    moments, reset) on a two-step state: a whole step is recorded by the
    stream call and launched together with the next one.  Reads after odd
    numbers of steps, a change of relaxation time and force, a boundary change
    (a new lid velocity), a stream without a collision and a tuning change all
    run the recorded step first; the populations equal those of one-step
    launches bit for bit, and plb_step's."""
    q, X, Y, Z = 19, 64, 12, 9
    ob, pb = bcs("z", lid=(0.02, 0.0, 0.0))
    _, pb2 = bcs("z", lid=(-0.01, 0.015, 0.0))
    mask = (np.random.default_rng(23).random((Z, Y, X)) < 0.08).astype(np.uint8)
    fl = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    F1, F2 = (1e-5, 0.0, 0.0), (2e-5, -1e-6, 0.0)
    with P.State(q, X, Y, Z, variant, P.FUSED) as a, P.State(q, X, Y, Z, variant, P.FUSED) as b, \
            P.State(q, X, Y, Z, variant, P.FUSED) as c:
        two_step(a, True)
        two_step(b, False)
        two_step(c, True)
        for st in (a, b, c):
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 4, 1e-3), pb)

        def staged(st, n, tau, F, bc, collide=True):
            for _ in range(n):
                if collide:
                    st.collide(tau, F)
                st.stream("push")
                st.apply_noslip(bc)
                st.apply_periodicity(bc)
                st.calc_moments()
                st.reset(bc)

        def same(tag):
            assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), tag
            assert np.array_equal(c.download_raw()[fl], b.download_raw()[fl]), tag

        for st in (a, b):
            staged(st, 7, 0.6, F1, pb)
        c.configure(0.6, F1, pb)
        c.step(7)
        same("7 steps")
        for st in (a, b):
            staged(st, 5, 0.6, F1, pb)
            staged(st, 4, 0.7, F2, pb)  # new relaxation time and force after an odd count
        c.step(5)
        c.configure(0.7, F2, pb)
        c.step(4)
        same("tau change")
        for st in (a, b):
            staged(st, 3, 0.7, F2, pb)
            # new lid velocity after an odd count: the staged calls hand the
            # boundary over after the stream, so the first of these steps
            # still streams with the old one
            staged(st, 3, 0.7, F2, pb2)
        c.step(4)
        c.configure(0.7, F2, pb2)
        c.step(2)
        same("boundary change")
        for st in (a, b):
            staged(st, 1, 0.7, F2, pb2)
            staged(st, 1, 0.7, F2, pb2, collide=False)  # a stream without a collision
            staged(st, 2, 0.7, F2, pb2)
        assert np.array_equal(a.download_raw()[fl], b.download_raw()[fl]), "stream only"
        staged(a, 3, 0.7, F2, pb2)
        a.set_tuning(capi.TUNE_TWO_STEP, 0)  # the recorded step runs before the change
        staged(a, 2, 0.7, F2, pb2)
        staged(b, 5, 0.7, F2, pb2)
        ra, ua = a.moments()
        rb, ub = b.moments()
        assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
        assert np.array_equal(a.global_moments(), b.global_moments())


@pytest.mark.parametrize("q,X,Y,Z,variant", [(19, 128, 100, 77, "f32s"), (19, 192, 41, 33, "f64"),
                                             (27, 384, 40, 29, "f32"), (19, 96, 57, 40, "f32s")])
def test_two_step_odd_shapes(q, X, Y, Z, variant):
    """Launch geometries the small oracle cases do not reach: (z-chunk,
    y-tile) pair counts that are not a multiple of 8 (a short last XCD run,
    and workgroups past it), ragged last tiles and z-chunks, 384-cell rows,
    32-cell tiles (x = 96), and workgroup i on tile i; against one-step
    launches, bit for bit, with solids and a moving lid."""
    ob, pb = bcs("z", lid=(0.02, 0.0, 0.01))
    mask = (np.random.default_rng(31).random((Z, Y, X)) < 0.1).astype(np.uint8)
    out = []
    for on, xcd in ((True, -1), (True, 0), (False, -1)):
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            two_step(st, on)
            st.set_tuning(capi.TUNE_XCD, xcd)
            assert st.scheme() == ("pull2" if on else "pull")
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0.01, 0.0, 0.0), (P.INIT_PERTURBED, 8, 1e-3), pb)
            st.configure(0.57, (1e-5, 0.0, 0.0), pb)
            st.step(9)
            out.append(st.download_raw())
    fluid = mask == 0
    fluid[0] = False
    fluid[-1] = False
    assert np.array_equal(out[0][fluid], out[2][fluid])
    assert np.array_equal(out[1][fluid], out[2][fluid])


def random_case(seed):
    """A seeded random lattice, boundary, storage variant and launch geometry."""
    r = np.random.default_rng(1000 + seed)
    q = int(r.choice([19, 19, 27]))
    variant = str(r.choice(VARIANTS[q]))
    X = int(r.choice([32, 64, 96, 128, 160, 192]))
    Y = int(r.integers(1, 24))
    Z = int(r.integers(1, 20))
    axes = "".join(a for a in "xyz" if r.random() < 0.4)
    if "y" in axes and Y < 3:
        axes = axes.replace("y", "")
    if "z" in axes and Z < 3:
        axes = axes.replace("z", "")
    faces = {"x": (EAST, WEST), "y": (NORTH, SOUTH), "z": (UP, DOWN)}
    t = [PERIODIC] * 6
    v = {}
    for a in axes:
        for f in faces[a]:
            t[f] = NOSLIP
            v[f] = tuple(float(x) for x in np.round(r.uniform(-0.03, 0.03, 3), 4)) if r.random() < 0.7 else (0.0, 0.0, 0.0)
    solids = float(r.choice([0.0, 0.05, 0.15]))
    mask = (r.random((Z, Y, X)) < solids).astype(np.uint8) if solids else None
    bx = 64 if X % 64 == 0 and r.random() < 0.7 and (q, variant) != (27, "f64") else 32
    wide = int(r.random() < 0.2)
    knobs = {capi.TUNE_TWO_ZCHUNK: int(r.choice([0, 0, 1, 3, 7])), capi.TUNE_TWO_TILE_X: bx,
             capi.TUNE_XCD: int(r.choice([-1, 0])), capi.TUNE_INDEX64: wide}
    steps = int(r.integers(3, 12))
    knobs[capi.TUNE_NT_STORE] = int(r.random() < 0.5)
    return q, variant, X, Y, Z, t, v, mask, knobs, steps


@pytest.mark.parametrize("seed", range(24))
def test_two_step_random_cases(orc, seed):
    """Seeded random lattices (rows of 32 to 192 cells, 1 to 23 rows, 1 to 19
    planes), walls on random axes with random wall velocities, interior solids,
    every storage variant, random z-chunk, tile width, tile mapping, index
    form and store kind: two steps per launch against one step per launch,
    bit for bit on the fluid cells, and the same moments everywhere."""
    q, variant, X, Y, Z, t, v, mask, knobs, steps = random_case(seed)
    print(f"\n[case] random {seed}: q={q} {variant} {X}x{Y}x{Z} walls={t} solids={mask is not None} "
          f"knobs={knobs} steps={steps}", flush=True)
    pb = P.boundary(t, v)
    out = []
    for on in (True, False):
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            st.set_tuning(capi.TUNE_TWO_STEP, 1 if on else 0)
            if on:
                for key, value in knobs.items():
                    st.set_tuning(key, value)
                assert st.scheme() == "pull2"
            st.set_boundary(pb, mask)
            st.initialize(1.0, (0.01, 0.0, -0.005), (P.INIT_PERTURBED, seed, 1e-3), pb)
            st.configure(0.58, (1e-5, 2e-6, -1e-6), pb)
            st.step(steps)
            out.append((st.download_raw(), st.moments()))
    fluid = orc.cor_flags(q, X, Y, Z, Boundary(type=t, vel=v), mask) == 0
    (ra, (rho_a, u_a)), (rb, (rho_b, u_b)) = out
    if variant == "f16s":
        assert np.array_equal(ra.view(np.uint16)[fluid], rb.view(np.uint16)[fluid])
    else:
        assert np.array_equal(ra[fluid], rb[fluid])
    assert np.array_equal(rho_a, rho_b) and np.array_equal(u_a, u_b)
