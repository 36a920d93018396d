"""The oracle's FP32 literal collision in the arithmetic of the reference's
AVX2 build (CollisionSRT.cpp:93-149, on by default in src/CMakeLists.txt:48-51),
pinned cell by cell against an independent evaluation of the same operation
sequence: float32 products and sums from numpy (each one correctly rounded),
and every fused multiply-add as the exact rational a*b + c rounded once to
float32.  The directions past the last full group of 8 are the scalar tail,
which must equal the oracle's scalar kernel."""
from fractions import Fraction

import numpy as np
import pytest

f32 = np.float32


def round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32, ties to even."""
    c = f32(float(x))
    best = None
    for cand in (np.nextafter(c, f32(-np.inf)), c, np.nextafter(c, f32(np.inf))):
        d = abs(Fraction(float(cand)) - x)
        even = (int(np.array(cand, f32).view(np.uint32)) & 1) == 0
        if best is None or d < best[0] or (d == best[0] and even):
            best = (d, cand)
    return f32(best[1])


def fma(a, b, c) -> np.float32:
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def avx2_cell(c, w, tau, F, rho, u, n, q):
    """_collision_kernel_avx2 for one cell, k < 8 * (q // 8)."""
    tau_r = f32(tau)
    om = f32(1.0 / float(tau_r))
    omm = f32(1) - om
    tri = tau_r / rho
    uu = [u[i] + f32(F[i]) * tri for i in range(3)]
    usq = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]
    kusq = f32(-1.5) * usq
    out = {}
    for k in range(8 * (q // 8)):
        cu = f32(0) if k == 0 else f32(c[k][0]) * uu[0] + f32(c[k][1]) * uu[1] + f32(c[k][2]) * uu[2]
        cusq = cu * cu
        neq = fma(f32(3), cu, f32(1))
        neq = fma(f32(4.5), cusq, neq)
        neq = neq + kusq
        neq = neq * rho
        neq = neq * f32(w[k])
        out[k] = fma(omm, n[k], om * neq)
    return out


@pytest.mark.parametrize("q", [9, 19, 27])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_avx2_collision_matches_exact_evaluation(orc, q, seed):
    rng = np.random.default_rng(seed)
    c, w, _, _ = orc.tables(q)
    tau, F = float(rng.uniform(0.52, 1.5)), tuple(float(v) for v in rng.uniform(-1e-3, 1e-3, 3))
    rho = f32(rng.uniform(0.9, 1.1))
    u = [f32(v) for v in rng.uniform(-0.05, 0.05, 3)]
    n = [f32(v) for v in rng.uniform(0.0, 0.4, q)]

    def run(avx2):
        st = orc.lit_state(q, 1, 1, 1, np.float32)
        st["rho"][1, 1, 1] = rho
        st["u"][1, 1, 1, :] = u
        st["n"][1, 1, 1, :] = n
        orc.lit_set_avx2(avx2)
        try:
            orc.lit_collide(st, tau, F)
        finally:
            orc.lit_set_avx2(False)
        return st["n"][1, 1, 1].copy()

    got, scalar = run(True), run(False)
    exp = avx2_cell(c, w, tau, F, rho, u, n, q)
    for k, v in exp.items():
        assert got[k] == v, (q, k, got[k], v)
    for k in range(8 * (q // 8), q):
        assert got[k] == scalar[k], (q, k)
    # the two kernels round differently but agree to a few units in the last place
    assert np.allclose(got, scalar, rtol=1e-6, atol=1e-7)
    assert not np.array_equal(got[:8 * (q // 8)], scalar[:8 * (q // 8)]) or q == 9
