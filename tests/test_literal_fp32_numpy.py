"""The oracle's FP32 literal step pinned against an independent numpy
evaluation written from the reference's source text: each stage is restated
here from the C++ types its operands have there (float fields, float weights,
int32 velocities, float members, double literals), not from the oracle.

Type facts the evaluation follows, with their source lines:
  * tau is float and omega = 1./tau is a double quotient stored in a float
    member (CollisionSRT.hpp:18-19, CollisionSRT.cpp:19-20);
  * u* = u + F * (tau / rho), all float (CollisionSRT.cpp:72-81);
  * usq and c.u* are float sums, left to right (CollisionSRT.cpp:84-91, 165);
  * neq = w_k * rho (float) times (1.0 + 3.0 cu + 4.5 cu cu - 1.5 usq)
    (double, from the double literals); the blend (1.0f - omega) * n is float,
    omega * neq double, their sum double, stored as float
    (CollisionSRT.cpp:172-175);
  * moments: float sums in ascending k, u scaled by the float 1.0f / rho
    (CalcMoments.hpp:22-37);
  * no-slip: n(c, k) - 2.0f * w_k * rho(y) * cu * cs2inv, float left to right,
    cu from float wall velocities, cs2inv = 1.0f / (1.0f / 3.0f)
    (Boundary.cpp:209-217, Boundary.cpp:27, LBModel.cpp:175-177); faces in the
    order east, west, north, south, up, down (Boundary.cpp:17-18, 101-111),
    each face in serial z, y, x, kp order;
  * periodic copies: x then y (Boundary.cpp:91-99, 223-281).
This is the arithmetic of the reference's FMA-free builds (its CI builds
Release with -DAVX2=OFF, .circleci/config.yml:11, and without -mfma no
contraction can happen); see the discrimination test at the end for the
promotion mistakes it tells apart.

The lattice tables come from the reference's own unit-test values
(tests/golden/reference_vectors.json), not from the oracle."""
import json
import os

import numpy as np
import pytest

f32, f64 = np.float32, np.float64
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
EAST, WEST, NORTH, SOUTH, UP, DOWN = range(6)
UNSET, PERIODIC, NOSLIP = 0, 1, 2


def tables(q):
    name = {9: "D2Q9", 19: "D3Q19", 27: "D3Q27"}[q]
    t = json.load(open(GOLDEN))["lbmodel"][name]
    # LBModel keeps the weights as float from double quotients such as 1./36.
    w = np.array([f32(eval(s, {"__builtins__": {}})) for s in t["w_text"]], f32)
    return np.array(t["c"], np.int32), w, np.array(t["rev"], np.int64)


class RefTextF32:
    """Halo'd AoS float fields (Field.hpp:91-93) and the staged operators."""

    def __init__(self, q, X, Y, Z):
        self.q, self.X, self.Y, self.Z = q, X, Y, Z
        self.c, self.w, self.rev = tables(q)

    # ---- CollisionSRT::_collision_kernel with _get_updated_u and _get_cu
    def collide(self, st, tau, F, mistake=None):
        n, rho, u = st["n"], st["rho"], st["u"]
        I = (slice(1, -1), slice(1, -1), slice(1, -1))
        tau_f = f32(tau)
        om = f32(f64(1.0) / f64(tau_f))
        omm = f32(1.0) - om
        rl = rho[I]
        tri = tau_f / rl
        uu = [u[I + (i,)] + f32(F[i]) * tri for i in range(3)]
        usq = (uu[0] * uu[0] + uu[1] * uu[1]) + uu[2] * uu[2]
        for k in range(self.q):
            if k == 0:
                cu = np.zeros_like(rl)
            else:
                ck = self.c[k]
                cu = (f32(ck[0]) * uu[0] + f32(ck[1]) * uu[1]) + f32(ck[2]) * uu[2]
            cud = cu.astype(f64)
            bracket = ((f64(1.0) + f64(3.0) * cud) + (f64(4.5) * cud) * cud) - f64(1.5) * usq.astype(f64)
            if mistake == "wr_in_double":
                wr = f64(self.w[k]) * rl.astype(f64)
            else:
                wr = (self.w[k] * rl).astype(f64)
            neq = wr * bracket
            a = omm * n[I + (k,)]
            if mistake == "blend_in_float":
                n[I + (k,)] = a + om * neq.astype(f32)
            elif mistake == "neq_in_float":
                b32 = ((f32(1.0) + f32(3.0) * cu) + (f32(4.5) * cu) * cu) - f32(1.5) * usq
                n[I + (k,)] = (a.astype(f64) + f64(om) * (self.w[k] * rl * b32).astype(f64)).astype(f32)
            else:
                n[I + (k,)] = (a.astype(f64) + f64(om) * neq).astype(f32)

    # ---- Streaming::_push_kernel, then std::swap(ntmp, n)
    def push(self, st):
        n, t = st["n"], st["ntmp"]
        X, Y, Z = self.X, self.Y, self.Z
        for k in range(self.q):
            cx, cy, cz = (int(v) for v in self.c[k])
            t[1 + cz:Z + 1 + cz, 1 + cy:Y + 1 + cy, 1 + cx:X + 1 + cx, k] = n[1:Z + 1, 1:Y + 1, 1:X + 1, k]
        st["n"], st["ntmp"] = t, n

    def face_cells(self, face):
        X, Y, Z = self.X, self.Y, self.Z
        lo, hi = [1, 1, 1], [X, Y, Z]  # inclusive, x y z
        axis = face // 2
        v = (X, Y, Z)[axis] if face % 2 == 0 else 1
        lo[axis] = hi[axis] = v
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    yield z, y, x

    # ---- Boundary::apply_noslip -> _apply_noslip, serial order
    def noslip(self, st, bc):
        n, rho = st["n"], st["rho"]
        cs2inv = f32(1.0) / (f32(1.0) / f32(3.0))
        for face in range(6):
            if bc["type"][face] != NOSLIP:
                continue
            ub = [f32(v) for v in bc["vel"][face]]
            for z, y, x in self.face_cells(face):
                for kp in range(1, self.q):
                    cx, cy, cz = (int(v) for v in self.c[kp])
                    nz, ny, nx = z + cz, y + cy, x + cx
                    k = int(self.rev[kp])
                    ck = self.c[k]
                    cu = (f32(ck[0]) * ub[0] + f32(ck[1]) * ub[1]) + f32(ck[2]) * ub[2]
                    term = ((f32(2.0) * self.w[k]) * rho[nz, ny, nx]) * cu * cs2inv
                    n[nz, ny, nx, kp] = n[z, y, x, k] - term

    # ---- Boundary::apply_periodicity: east-west, then north-south
    def periodic(self, st, bc):
        n = st["n"]
        X, Y = self.X, self.Y
        if bc["type"][EAST] == PERIODIC and bc["type"][WEST] == PERIODIC:
            for k in range(1, self.q):
                if self.c[k][0] > 0:
                    n[1:-1, 1:-1, 1, k] = n[1:-1, 1:-1, X + 1, k]
                if self.c[k][0] < 0:
                    n[1:-1, 1:-1, X, k] = n[1:-1, 1:-1, 0, k]
        if bc["type"][NORTH] == PERIODIC and bc["type"][SOUTH] == PERIODIC:
            for k in range(1, self.q):
                if self.c[k][1] > 0:
                    n[1:-1, 1, 1:-1, k] = n[1:-1, Y + 1, 1:-1, k]
                if self.c[k][1] < 0:
                    n[1:-1, Y, 1:-1, k] = n[1:-1, 0, 1:-1, k]

    # ---- CalcMoments::_get_local_moments
    def moments(self, st):
        n = st["n"]
        I = (slice(1, -1), slice(1, -1), slice(1, -1))
        rl = np.zeros(n[I + (0,)].shape, f32)
        ul = [np.zeros_like(rl) for _ in range(3)]
        for k in range(self.q):
            nk = n[I + (k,)]
            rl = rl + nk
            for i in range(3):
                ul[i] = ul[i] + nk * f32(self.c[k][i])
        rhoinv = f32(1.0) / rl
        st["rho"][I] = rl
        for i in range(3):
            st["u"][I + (i,)] = ul[i] * rhoinv

    # ---- Boundary::reset = reset_velocity, reset_density
    def reset(self, st, bc):
        for face in range(6):
            if face in bc["vel"]:
                for z, y, x in self.face_cells(face):
                    st["u"][z, y, x, :] = [f32(v) for v in bc["vel"][face]]
        for face in range(6):
            if bc["type"][face] == NOSLIP:
                for z, y, x in self.face_cells(face):
                    st["rho"][z, y, x] = f32(bc["rho_s"])


CASES = [  # q, X, Y, Z, face types, wall velocities
    (9, 8, 1, 6, [PERIODIC, PERIODIC, UNSET, UNSET, NOSLIP, NOSLIP], {UP: (0.05, 0.0, 0.01), DOWN: (0.0, 0.0, 0.0)}),
    (19, 5, 4, 3, [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {UP: (0.05, 0.01, 0.0), DOWN: (0.0, -0.002, 0.0)}),
    (19, 4, 5, 4, [NOSLIP] * 6, {f: (0.01 * (f + 1), -0.003 * f, 0.002) for f in range(6)}),
    (27, 5, 4, 3, [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {UP: (0.04, 0.0, 0.003), DOWN: (0.0, 0.0, 0.0)}),
    (27, 4, 3, 5, [NOSLIP, NOSLIP, PERIODIC, PERIODIC, NOSLIP, NOSLIP],
     {EAST: (0.0, 0.01, 0.0), WEST: (0.0, 0.0, 0.0), UP: (0.03, 0.0, 0.0), DOWN: (0.0, 0.0, -0.01)}),
]


def random_state(q, X, Y, Z, seed):
    rng = np.random.default_rng(seed)
    shp = (Z + 2, Y + 2, X + 2)
    _, w, _ = tables(q)
    return {
        "rho": rng.uniform(0.95, 1.05, shp).astype(f32),
        "u": rng.uniform(-0.05, 0.05, shp + (3,)).astype(f32),
        "n": (w * rng.uniform(0.8, 1.2, shp + (q,))).astype(f32),
        "ntmp": (w * rng.uniform(0.8, 1.2, shp + (q,))).astype(f32),
    }


def orc_boundary(types, vel, rho_s):
    from oracle.pyoracle import Boundary
    return Boundary(type=list(types), vel=dict(vel), rho_s=rho_s)


def copy_into(orc, st, q, X, Y, Z):
    o = orc.lit_state(q, X, Y, Z, np.float32)
    for key in ("rho", "u", "n", "ntmp"):
        o[key][...] = st[key]
    return o


def assert_same(o, st, where):
    for key in ("rho", "u", "n", "ntmp"):
        a, b = o[key], st[key]
        if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
            bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
            raise AssertionError(f"{where}: {key} differs at {len(bad)} entries, first {bad[:3].tolist()}: "
                                 f"oracle {a[tuple(bad[0])]!r} numpy {b[tuple(bad[0])]!r}")


@pytest.mark.parametrize("q,X,Y,Z,types,vel", CASES)
def test_each_fp32_stage_matches_the_reference_text(orc, q, X, Y, Z, types, vel):
    """Every stage from the same random state: oracle == numpy, bit for bit."""
    ref = RefTextF32(q, X, Y, Z)
    bc = {"type": types, "vel": vel, "rho_s": 1.02}
    obc = orc_boundary(types, vel, 1.02)
    tau, F = 0.57, (2e-4, -1e-4 if q != 9 else 0.0, 5e-5)
    for i, stage in enumerate(("collide", "push", "noslip", "periodic", "moments", "reset")):
        st = random_state(q, X, Y, Z, seed=1000 * q + 10 * X + i)
        o = copy_into(orc, st, q, X, Y, Z)
        if stage == "collide":
            ref.collide(st, tau, F)
            orc.lit_collide(o, tau, F)
        elif stage == "push":
            ref.push(st)
            orc.lit_push(o)
        elif stage == "noslip":
            ref.noslip(st, bc)
            assert orc.lit_noslip(o, obc) == 0
        elif stage == "periodic":
            ref.periodic(st, bc)
            orc.lit_periodic(o, obc)
        elif stage == "moments":
            ref.moments(st)
            orc.lit_moments(o)
        else:
            ref.reset(st, bc)
            orc.lit_reset(o, obc)
        assert_same(o, st, stage)


@pytest.mark.parametrize("q,X,Y,Z,types,vel", CASES)
def test_fp32_steps_match_the_reference_text(orc, q, X, Y, Z, types, vel):
    """Four whole steps of PLBPD::march_in_time (PLBPD.cpp:24-29) from the
    initial state of Domain::initialize plus Boundary::reset."""
    ref = RefTextF32(q, X, Y, Z)
    bc = {"type": types, "vel": vel, "rho_s": 1.0}
    obc = orc_boundary(types, vel, 1.0)
    tau, F = 0.61, (1e-4, 3e-5 if q != 9 else 0.0, -2e-5)
    o = orc.lit_state(q, X, Y, Z, np.float32)
    orc.lit_init(o, 1.0, (0.01, 0.0, 0.0), (1, 77, 1e-2), obc)
    st = {k: o[k].copy() for k in ("rho", "u", "n", "ntmp")}
    for step in range(4):
        orc.lit_collide(o, tau, F)
        orc.lit_push(o)
        assert orc.lit_noslip(o, obc) == 0
        orc.lit_periodic(o, obc)
        orc.lit_moments(o)
        orc.lit_reset(o, obc)
        ref.collide(st, tau, F)
        ref.push(st)
        ref.noslip(st, bc)
        ref.periodic(st, bc)
        ref.moments(st)
        ref.reset(st, bc)
        assert_same(o, st, f"step {step}")


@pytest.mark.parametrize("mistake", ["wr_in_double", "blend_in_float", "neq_in_float"])
def test_promotion_mistakes_are_told_apart(orc, mistake):
    """Each plausible promotion mistake in the collision (w*rho in double,
    the blend in float, the bracket in float) gives another float somewhere in
    one D3Q19 collision of a random state, so the stage test above fails on it."""
    q, X, Y, Z = 19, 6, 5, 4
    ref = RefTextF32(q, X, Y, Z)
    st = random_state(q, X, Y, Z, seed=5)
    o = copy_into(orc, st, q, X, Y, Z)
    ref.collide(st, 0.57, (2e-4, -1e-4, 5e-5), mistake=mistake)
    orc.lit_collide(o, 0.57, (2e-4, -1e-4, 5e-5))
    assert not np.array_equal(o["n"].view(np.uint32), st["n"].view(np.uint32))
