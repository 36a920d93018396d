"""Pins of the oracle's corrected (fused) step beyond the FP64 literal chain:
the storage variants (FP32, FP64 storage with FP32 arithmetic, FP32 and FP16
shifted deviations) and interior solid voxels.

The FP64 corrected step is pinned against the literal step elsewhere (closed
box bit for bit, seams only otherwise, analytic flows); here every reduced
variant is held to the FP64 corrected run, the shifted forms to their exact
stored values, and solids to wall faces, mass conservation and mirror
symmetry, so that a wrong shift, scale, sign or reverse index fails."""
import numpy as np
import pytest

from oracle.pyoracle import Boundary, PERIODIC, channel_bc

Q19_TOL = {"f32s": 2e-6, "f32": 2e-3, "f64a32": 2e-3, "f16s": 2e-3}


def _run(orc, variant, q, dims, tau, F, bc, steps, init, solid=None):
    X, Y, Z = dims
    st = orc.cor_state(variant, q, X, Y, Z)
    orc.cor_init(st, 1.0, (0, 0, 0), init)
    orc.cor_run(st, tau, F, bc, steps, solid)
    rho, u = orc.cor_moments(st, bc, solid)
    return st, rho, u


def _solids(dims, frac, seed):
    X, Y, Z = dims
    return (np.random.default_rng(seed).random((Z, Y, X)) < frac).astype(np.uint8)


@pytest.mark.parametrize("variant", ["f32s", "f32", "f64a32", "f16s"])
def test_variant_tracks_fp64(orc, variant):
    """A channel with solids, 300 steps: each reduced-precision variant stays
    within its stated velocity L2 bound of the FP64 run; shifted FP32 is far
    closer than plain FP32 (the point of the shift)."""
    q, dims = 19, (20, 12, 10)
    bc = channel_bc()
    solid = _solids(dims, 0.1, 7)
    args = (q, dims, 0.6, (2e-5, 0, 0), bc, 300, (1, 5, 1e-3), solid)
    _, rho64, u64 = _run(orc, "f64", *args)
    _, rho, u = _run(orc, variant, *args)
    fluid = orc.cor_flags(q, *dims, bc, solid) == 0
    du = u[fluid] - u64[fluid]
    err = np.sqrt((du ** 2).sum() / (u64[fluid] ** 2).sum())
    assert err < Q19_TOL[variant], err
    assert err > 0.0  # the variant really is a different arithmetic
    assert np.abs(rho[fluid] - rho64[fluid]).max() < 1e-4


def test_fp64_storage_fp32_arithmetic_equals_fp32_storage(orc):
    """FP64 storage with FP32 arithmetic rounds every stored value to float,
    so its trajectory is the FP32-storage trajectory exactly on fluid cells
    (wall cells keep their unrounded initial values, which nothing reads)."""
    q, dims = 19, (12, 10, 8)
    bc = channel_bc(walls_vel_up=(0.02, 0.0, 0.0))
    args = (q, dims, 0.7, (1e-5, 0, 0), bc, 40, (1, 11, 1e-3))
    a, _, _ = _run(orc, "f64a32", *args)
    b, _, _ = _run(orc, "f32", *args)
    fluid = orc.cor_flags(q, *dims, bc) == 0
    assert np.array_equal(a["f"][fluid], b["f"].astype(np.float64)[fluid])


@pytest.mark.parametrize("q", [19, 27])
def test_shifted_rest_state_is_zero(orc, q):
    """At rest (rho = 1, u = 0) the shifted forms store exactly zero, and
    adding the weights back returns the weights."""
    X, Y, Z = 4, 3, 5
    w = orc.tables(q)[1]
    for variant in ("f32s", "f16s"):
        st = orc.cor_state(variant, q, X, Y, Z)
        orc.cor_init(st, 1.0, (0, 0, 0), (0, 0, 0.0))
        assert not st["f"].any()
        pop = orc.cor_populations(st)
        assert np.array_equal(pop, np.broadcast_to(w.astype(np.float32).astype(np.float64),
                                                   pop.shape)) or np.allclose(pop, w, rtol=0, atol=1e-8)


def test_shifted_stored_values(orc):
    """f32s stores float(f - w); f16s stores half(32768 (f - w)).  Checked
    against FP64 equilibria of the same perturbed initial state."""
    q, X, Y, Z = 19, 6, 5, 4
    w = orc.tables(q)[1].astype(np.float64)
    ref = orc.cor_state("f64", q, X, Y, Z)
    orc.cor_init(ref, 1.0, (0, 0, 0), (1, 3, 2e-2))
    g64 = ref["f"] - w
    s32 = orc.cor_state("f32s", q, X, Y, Z)
    orc.cor_init(s32, 1.0, (0, 0, 0), (1, 3, 2e-2))
    assert np.allclose(s32["f"].astype(np.float64), g64, rtol=2 ** -23, atol=1e-12)
    s16 = orc.cor_state("f16s", q, X, Y, Z)
    orc.cor_init(s16, 1.0, (0, 0, 0), (1, 3, 2e-2))
    h = s16["f"].view(np.float16).astype(np.float64) / 32768.0
    assert np.all(np.abs(h - g64) <= np.abs(g64) * 2 ** -11 + 2 ** -39)
    assert np.abs(g64).max() > 1e-3  # the perturbation really is there


def test_solid_planes_equal_wall_faces(orc):
    """A z-periodic box closed by one solid plane is the no-slip channel with
    one layer fewer: solid-voxel bounce-back equals the wall-face path."""
    q, X, Y = 19, 10, 8
    F, tau, steps, init = (3e-5, 1e-5, 0), 0.65, 60, (0, 0, 0.0)
    chan = channel_bc()
    _, rho_c, u_c = _run(orc, "f64", q, (X, Y, 10), tau, F, chan, steps, init)
    box = Boundary(type=[PERIODIC] * 6, vel={})
    solid = np.zeros((9, Y, X), np.uint8)
    solid[0] = 1
    st_b, rho_b, u_b = _run(orc, "f64", q, (X, Y, 9), tau, F, box, steps, init, solid)
    assert np.abs(u_c[1:9]).max() > 1e-4
    assert np.array_equal(rho_c[1:9], rho_b[1:9])
    assert np.array_equal(u_c[1:9], u_b[1:9])


@pytest.mark.parametrize("variant", ["f64", "f32s"])
def test_solids_conserve_mass(orc, variant):
    """Stationary bounce-back moves populations, never creates them: total
    mass over fluid cells is conserved (FP64 exactly up to summation order,
    shifted FP32 to float rounding of the deviations)."""
    q, dims = 19, (14, 12, 10)
    box = Boundary(type=[PERIODIC] * 6, vel={})
    solid = _solids(dims, 0.2, 3)
    st = orc.cor_state(variant, q, *dims)
    orc.cor_init(st, 1.0, (0, 0, 0), (1, 2, 1e-2))
    fluid = orc.cor_flags(q, *dims, box, solid) == 0
    m0 = orc.cor_populations(st)[fluid].sum()
    orc.cor_run(st, 0.8, (0, 0, 0), box, 50, solid)
    m1 = orc.cor_populations(st)[fluid].sum()
    tol = 1e-12 if variant == "f64" else 1e-8
    assert abs(m1 - m0) / m0 < tol


def test_solids_mirror_symmetry(orc):
    """A mask symmetric under y -> Y-1-y with force along x gives a flow whose
    u_x and rho are mirror-symmetric and u_y antisymmetric (a wrong reverse
    index or neighbour offset breaks this)."""
    q, dims = 19, (12, 10, 8)
    X, Y, Z = dims
    half = _solids((X, Y // 2, Z), 0.15, 9)
    solid = np.concatenate([half, half[:, ::-1]], axis=1)
    bc = channel_bc()
    _, rho, u = _run(orc, "f64", q, dims, 0.7, (4e-5, 0, 0), bc, 80, (0, 0, 0.0), solid)
    fluid = orc.cor_flags(q, *dims, bc, solid) == 0
    m = fluid & fluid[:, ::-1]
    assert np.abs(u[..., 1][m]).max() > 1e-7
    assert np.allclose(rho[m], rho[:, ::-1][m], rtol=1e-13, atol=0)
    scale = 1e-11 * np.abs(u[m]).max()
    assert np.abs(u[..., 0][m] - u[:, ::-1, :, 0][m]).max() < scale
    assert np.abs(u[..., 1][m] + u[:, ::-1, :, 1][m]).max() < scale
