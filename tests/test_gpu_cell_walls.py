"""The cell pull kernel without interior solids.  A wavefront none of whose
cells borders a wall face takes every population from x - c_k without the
bounce-back select; the others select per cell.  Rows of 256 cells (four
wavefronts: with walls on x the first and last border a wall, the middle two
do not), walls on x, y and z in turn, a moving lid where z has walls, with
32- and 64-bit indices, against the oracle's corrected step bit for bit."""
import numpy as np
import pytest

import plbpd_amd as P
from plbpd_amd import capi
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, EAST, WEST, NORTH, SOUTH, UP, DOWN

pytestmark = pytest.mark.gpu


def walls(axes):
    types, vel = [PERIODIC] * 6, {}
    for f, ax in ((EAST, "x"), (WEST, "x"), (NORTH, "y"), (SOUTH, "y"), (UP, "z"), (DOWN, "z")):
        if ax in axes:
            types[f] = NOSLIP
            vel[f] = (0.0, 0.0, 0.0)
    if "z" in axes:
        vel[UP] = (0.02, 0.0, 0.01)
    return Boundary(type=types, vel=vel), P.boundary(types, vel)


@pytest.mark.parametrize("index64", [0, 1])
@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("axes", ["x", "yz", "xyz", ""])
def test_cell_kernel_walls_only(orc, axes, q, index64):
    X, Y, Z, steps = 256, 5, 4, 7
    ob, pb = walls(axes)
    tau, F = 0.58, (1e-5, -3e-6, 2e-6)
    init = (P.INIT_PERTURBED, 31, 1e-3)
    fluid = orc.cor_flags(q, X, Y, Z, ob, None) == 0
    for variant in ("f64", "f32", "f32s", "f16s"):
        print(f"\n[case] cell kernel walls={axes or 'none'} q={q} {variant} index64={index64}", flush=True)
        o = orc.cor_state(variant, q, X, Y, Z)
        orc.cor_init(o, 1.0, (0.01, 0.0, 0.0), init)
        orc.cor_run(o, tau, F, ob, steps, None)
        with P.State(q, X, Y, Z, variant, P.FUSED) as st:
            st.set_tuning(capi.TUNE_TWO_STEP, 0)
            st.set_tuning(capi.TUNE_VECTOR, 0)
            st.set_tuning(capi.TUNE_INDEX64, index64)
            st.set_boundary(pb, None)
            st.initialize(1.0, (0.01, 0.0, 0.0), init, pb)
            st.configure(tau, F, pb)
            assert st.scheme() == "pull"
            st.step(steps)
            raw = st.download_raw()
        if variant == "f16s":
            same = np.array_equal(raw.view(np.uint16)[fluid], o["f"].view(np.uint16)[fluid])
        else:
            same = np.array_equal(raw[fluid], o["f"][fluid])
        assert same, (axes, q, variant, index64)
