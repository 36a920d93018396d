"""Moving walls on two and three axes at once, a different velocity on every
face: the wall cells along the edges and corners lie on two or three faces,
and take the velocity of the last face in the reference's order (east, west,
north, south, up, down), as the oracle's flags do.  Every fused kernel that
writes moving-wall terms, against the oracle's corrected step, bit for bit:
two steps per launch, the row pull kernel, the cell pull kernel, the push
kernel and the one-array (AA) kernels."""
import numpy as np
import pytest

import plbpd_amd as P
from plbpd_amd import capi
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, EAST, WEST, NORTH, SOUTH, UP, DOWN

pytestmark = pytest.mark.gpu

FACES = {
    EAST: (0.0, 0.02, 0.01),
    WEST: (0.01, 0.0, -0.02),
    NORTH: (0.02, 0.0, -0.01),
    SOUTH: (0.0, 0.01, 0.015),
    UP: (0.03, -0.01, 0.0),
    DOWN: (-0.01, 0.0, 0.02),
}


def bcs(axes):
    t = [PERIODIC] * 6
    v = {}
    for f, on in ((EAST, "x"), (WEST, "x"), (NORTH, "y"), (SOUTH, "y"), (UP, "z"), (DOWN, "z")):
        if on in axes:
            t[f] = NOSLIP
            v[f] = FACES[f]
    return Boundary(type=t, vel=v), P.boundary(t, v)


# how each kernel is chosen: (mode, tuning settings)
KERNELS = {
    "two-step": (P.FUSED, {capi.TUNE_TWO_STEP: 1}),
    "row": (P.FUSED, {capi.TUNE_TWO_STEP: 0, capi.TUNE_VECTOR: 1}),
    "cell": (P.FUSED, {capi.TUNE_TWO_STEP: 0, capi.TUNE_VECTOR: 0}),
    "push": (P.FUSED, {capi.TUNE_TWO_STEP: 0, capi.TUNE_SCHEME: capi.SCHEME_PUSH}),
    "aa": (P.FUSED_AA, {}),
}


@pytest.mark.parametrize("axes", ["yz", "xz", "xyz"])
@pytest.mark.parametrize("q,variants", [(19, ["f64", "f32s", "f16s"]), (27, ["f32s"])])
def test_moving_walls_on_several_axes(orc, axes, q, variants):
    X, Y, Z, steps = 64, 10, 9, 9
    ob, pb = bcs(axes)
    tau, F = 0.6, (1e-5, -2e-6, 1e-6)
    mask = (np.random.default_rng(23).random((Z, Y, X)) < 0.08).astype(np.uint8)
    init = (P.INIT_PERTURBED, 12, 1e-3)
    fluid = orc.cor_flags(q, X, Y, Z, ob, mask) == 0
    for variant in variants:
        o = orc.cor_state(variant, q, X, Y, Z)
        orc.cor_init(o, 1.0, (0.0, 0.0, 0.0), init)
        orc.cor_run(o, tau, F, ob, steps, mask)
        orho, ou = orc.cor_moments(o, ob, mask)
        for name, (mode, tuning) in KERNELS.items():
            if name == "two-step" and q == 27 and variant == "f64":
                continue
            print(f"\n[case] walls {axes} q={q} {variant} {name}", flush=True)
            with P.State(q, X, Y, Z, variant, mode) as st:
                for key, value in tuning.items():
                    st.set_tuning(key, value)
                if name == "two-step":
                    assert st.scheme() == "pull2"
                st.set_boundary(pb, mask)
                st.initialize(1.0, (0.0, 0.0, 0.0), init, pb)
                st.configure(tau, F, pb)
                st.step(steps)
                raw = st.download_raw()
                rho, u = st.moments()
            if variant == "f16s":
                same = np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid])
            else:
                same = np.array_equal(raw[fluid], o["f"][fluid])
            assert same, (axes, q, variant, name)
            assert np.array_equal(rho, orho) and np.array_equal(u, ou), (axes, q, variant, name)
