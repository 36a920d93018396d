"""The literal collide+push kernel stores its in-place post-collision values
(the reference's ntmp after the swap, CollisionSRT.cpp:172-175 then
Streaming.cpp:82, 93) on the outer interior layer only, and rebuilds the rest
on demand as ntmp(c, k) = n(c + c_k, k).  This checks that argument on the
CPU oracle: after every collide+push the values off the outer layer are
replaced by NaN, and the run must be unchanged -- every later stage of
PLBPD::march_in_time (PLBPD.cpp:24-29) and the next push must never read
them -- while the rebuild rule must reproduce the reference's ntmp exactly at
every stage up to the next collision."""
import numpy as np
import pytest

from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, UNSET, UP, DOWN, EAST, WEST, NORTH, SOUTH


def deep(st):
    """Mask of interior cells off the outer layer (halo'd shape)."""
    Z, Y, X = st["Z"], st["Y"], st["X"]
    m = np.zeros((Z + 2, Y + 2, X + 2), bool)
    m[2:Z, 2:Y, 2:X] = True
    return m


def rebuilt_ntmp(orc, st):
    """ntmp with the cells off the outer layer taken from n(c + c_k, k)."""
    c = orc.tables(st["q"])[0]
    t = st["ntmp"].copy()
    Z, Y, X = st["Z"], st["Y"], st["X"]
    for k in range(st["q"]):
        cx, cy, cz = (int(v) for v in c[k])
        t[2:Z, 2:Y, 2:X, k] = st["n"][2 + cz:Z + cz, 2 + cy:Y + cy, 2 + cx:X + cx, k]
    return t


def same(a, b):
    return np.array_equal(a, b, equal_nan=False)


CASES = [  # q, X, Y, Z, face types, wall velocities
    (19, 8, 6, 5, [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {UP: (0.05, 0.01, 0.0), DOWN: (0.0, 0.0, 0.0)}),
    (19, 6, 5, 7, [NOSLIP] * 6, {f: (0.01 * f, -0.004, 0.002 * f) for f in range(6)}),
    (27, 7, 5, 6, [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {UP: (0.04, 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)}),
    (27, 5, 6, 5, [NOSLIP, NOSLIP, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {EAST: (0.0, 0.02, 0.0), WEST: (0.0, 0.0, 0.01), UP: (0.03, 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)}),
    (9, 8, 1, 6, [PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP], {UP: (0.05, 0.0, 0.0), DOWN: (0, 0, 0)}),
]


@pytest.mark.parametrize("q,X,Y,Z,types,vel", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_inplace_values_off_the_outer_layer_are_never_read(orc, q, X, Y, Z, types, vel, dtype):
    bc = Boundary(type=types, vel=vel, rho_s=1.0)
    tau, F = 0.63, (1e-4, 2e-5 if q != 9 else 0.0, -1e-5)
    ref = orc.lit_state(q, X, Y, Z, dtype)
    orc.lit_init(ref, 1.0, (0.01, 0.0, 0.0), (1, 5, 1e-2), bc)
    eli = orc.lit_state(q, X, Y, Z, dtype)
    orc.lit_init(eli, 1.0, (0.01, 0.0, 0.0), (1, 5, 1e-2), bc)
    m = deep(ref)
    stages = [("noslip", lambda s: orc.lit_noslip(s, bc)), ("periodic", lambda s: orc.lit_periodic(s, bc)),
              ("moments", orc.lit_moments), ("reset", lambda s: orc.lit_reset(s, bc))]
    for step in range(6):
        for s in (ref, eli):
            orc.lit_collide(s, tau, F)
            orc.lit_push(s)
        eli["ntmp"][m] = np.nan  # the in-place stores the kernel leaves out
        checks = [("push", None)] + stages
        for name, fn in checks:
            if fn is not None:
                for s in (ref, eli):
                    fn(s)
            where = f"step {step} after {name}"
            for key in ("rho", "u", "n"):
                assert same(eli[key], ref[key]), (where, key)
            assert same(eli["ntmp"][~m], ref["ntmp"][~m]), where
            assert same(rebuilt_ntmp(orc, eli), ref["ntmp"]), where
    if q == 9:  # Y = 1: every cell is on the outer layer and nothing is left out
        assert not m.any()
    else:
        assert m.sum() > 0
