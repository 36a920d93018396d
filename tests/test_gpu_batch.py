"""Batched domains (one launch per step for many channel domains) against the
CPU oracle's corrected step, domain by domain, bit for bit."""
import numpy as np
import pytest

import plbpd_amd as P
from oracle.pyoracle import Boundary, NOSLIP, PERIODIC, UP, DOWN

pytestmark = pytest.mark.gpu


def domains(n, seed, lo=6, hi=22):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y, z = (int(v) for v in rng.integers(lo, hi, 3))
        out.append({"x": x, "y": y, "z": z, "tau": float(rng.uniform(0.55, 1.5)),
                    "force": (float(rng.uniform(1e-7, 1e-5)), 0.0, 0.0)})
    return out


def channel(lid=(0.0, 0.0, 0.0)):
    types = [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP]
    vel = {UP: lid, DOWN: (0.0, 0.0, 0.0)}
    return Boundary(type=types, vel=vel), P.boundary(types, vel)


def oracle_moments(orc, variant, dm, d, ob, steps, seed, amp):
    o = orc.cor_state(variant, 19, dm["x"], dm["y"], dm["z"])
    orc.cor_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, seed + d, amp))
    orc.cor_run(o, dm["tau"], dm["force"], ob, steps)
    return orc.cor_moments(o, ob)


@pytest.mark.parametrize("variant", ["f64", "f32s", "f16s"])
@pytest.mark.parametrize("mode", ["plain", "graph", "per_domain", "persistent", "plain_xcd", "persistent_xcd"])
@pytest.mark.parametrize("scheme", ["pull", "push"])
def test_batch_matches_oracle(orc, variant, mode, scheme):
    print(f"\n[case] batch {variant} {mode} {scheme}", flush=True)
    doms = domains(7, 11)
    ob, pb = channel()
    steps, seed, amp = 12, 40, 2e-3
    b = P.Batch(19, variant, doms, pb)
    b.set_scheme(2 if scheme == "pull" else 1)
    if mode.startswith("persistent"):
        b.set_tiling(128, 3)
    elif mode == "per_domain":
        b.set_tiling(64, 0)
    if mode.endswith("_xcd"):
        b.set_mapping(1)
    b.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, seed, amp))
    if mode == "graph":
        b.step(steps, steps_per_graph=4)
    elif mode == "per_domain":
        b.step_per_domain(steps)
    else:
        b.step(steps)
    b.synchronize()
    for d, dm in enumerate(doms):
        rho, u = b.moments(d)
        orho, ou = oracle_moments(orc, variant, dm, d, ob, steps, seed, amp)
        assert np.array_equal(rho, orho), f"domain {d} rho"
        assert np.array_equal(u, ou), f"domain {d} u"
    b.close()


@pytest.mark.parametrize("scheme", ["pull", "push"])
@pytest.mark.parametrize("variant", ["f64", "f16s"])
def test_batch_moving_lid(orc, scheme, variant):
    doms = domains(4, 3)
    ob, pb = channel(lid=(0.02, 0.0, 0.0))
    b = P.Batch(19, variant, doms, pb)
    b.set_scheme(2 if scheme == "pull" else 1)
    b.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, 5, 1e-3))
    b.step(4)
    b.moments(0)  # a read in the middle of the run (pull: finishes the pending streaming)
    b.step(3, steps_per_graph=2)
    b.step(3)
    for d, dm in enumerate(doms):
        rho, u = b.moments(d)
        orho, ou = oracle_moments(orc, variant, dm, d, ob, 10, 5, 1e-3)
        assert np.array_equal(rho, orho) and np.array_equal(u, ou)
    b.close()


@pytest.mark.parametrize("q", [9, 27])
@pytest.mark.parametrize("variant", ["f64", "f32s"])
@pytest.mark.parametrize("scheme", ["pull", "push"])
@pytest.mark.parametrize("tile", [64, 256])
def test_batch_other_lattices_and_ragged_domains(orc, q, variant, scheme, tile):
    """D2Q9 and D3Q27 batches with domains that end mid-tile, one cell wide,
    three layers thin (two walls and one fluid layer), and filling whole
    tiles, next to each other in one pool."""
    print(f"\n[case] batch q={q} {variant} {scheme} tile {tile}", flush=True)
    if q == 9:
        types = [PERIODIC, PERIODIC, 0, 0, NOSLIP, NOSLIP]
        shapes = [(1, 1, 3), (64, 1, 4), (37, 1, 9), (5, 1, 3), (130, 1, 6)]
    else:
        types = [PERIODIC, PERIODIC, PERIODIC, PERIODIC, NOSLIP, NOSLIP]
        shapes = [(1, 1, 3), (4, 4, 4), (7, 3, 5), (1, 5, 6), (16, 8, 3), (9, 11, 7)]
    vel = {UP: (0.0, 0.0, 0.0), DOWN: (0.0, 0.0, 0.0)}
    ob, pb = Boundary(type=types, vel=vel), P.boundary(types, vel)
    rng = np.random.default_rng(q)
    doms = [{"x": x, "y": y, "z": z, "tau": float(rng.uniform(0.6, 1.2)),
             "force": (float(rng.uniform(1e-6, 1e-5)), 0.0, 0.0)} for x, y, z in shapes]
    steps, seed, amp = 9, 70, 3e-3
    b = P.Batch(q, variant, doms, pb)
    b.set_scheme(2 if scheme == "pull" else 1)
    b.set_tiling(tile, 0)
    b.initialize(1.0, (0, 0, 0), (P.INIT_PERTURBED, seed, amp))
    b.step(steps)
    b.synchronize()
    for d, dm in enumerate(doms):
        o = orc.cor_state(variant, q, dm["x"], dm["y"], dm["z"])
        orc.cor_init(o, 1.0, (0, 0, 0), (P.INIT_PERTURBED, seed + d, amp))
        orc.cor_run(o, dm["tau"], dm["force"], ob, steps)
        orho, ou = orc.cor_moments(o, ob)
        rho, u = b.moments(d)
        assert np.array_equal(rho, orho), f"domain {d} {dm['x']}x{dm['y']}x{dm['z']} rho"
        assert np.array_equal(u, ou), f"domain {d} {dm['x']}x{dm['y']}x{dm['z']} u"
    b.close()


def test_batch_autotune_and_cells():
    doms = domains(16, 9, 16, 40)
    _, pb = channel()
    b = P.Batch(19, "f32s", doms, pb)
    b.initialize()
    res = b.autotune(steps=4)
    assert len(res) > 0 and np.all(res[:, 3] > 0)
    tile, grid = b.tiling()
    best = res[np.argmin(res[:, 3])]
    assert (tile, grid, b.mapping()) == (int(best[0]), int(best[1]), int(best[2]))
    total, fluid = b.cells()
    assert total == sum(d["x"] * d["y"] * d["z"] for d in doms)
    assert fluid == sum(d["x"] * d["y"] * (d["z"] - 2) for d in doms)
    b.close()


def test_autotune_advances_the_documented_steps():
    """plb_batch_autotune runs real steps: every candidate advances the batch
    by 2 + (steps rounded up to even).  After it, the batch equals a twin
    stepped that many times per reported candidate, bit for bit."""
    print("\n[case] batch autotune advance", flush=True)
    doms = domains(5, 3)
    _, pb = channel()
    init = (P.INIT_PERTURBED, 9, 1e-3)
    a = P.Batch(19, "f32s", doms, pb)
    b = P.Batch(19, "f32s", doms, pb)
    for x in (a, b):
        x.initialize(1.0, (0, 0, 0), init)
    steps = 3
    res = a.autotune(steps=steps, max_results=128)
    assert len(res) == 60
    b.step(len(res) * (2 + steps + (steps & 1)))
    for d in range(len(doms)):
        ra, ua = a.moments(d)
        rb, ub = b.moments(d)
        assert np.array_equal(ra, rb) and np.array_equal(ua, ub), d
    a.close()
    b.close()


def test_batch_f32s_velocity_error_bound():
    """FP32-shifted batched domains against an FP64 batch: every domain's
    velocity L2 error stays at or below 2^-24 per step.  The largest errors
    are in domains whose bulk still accelerates uniformly (low viscosity,
    tall channels: u = g t there).  All bulk cells then hold the same values,
    so each step's FP32 rounding has the same sign everywhere and adds up
    with the step count rather than averaging out; domains that reached their
    steady profile stop gaining error."""
    from plbpd_amd.reports import c5_domains
    print("\n[case] batch f32s error bound", flush=True)
    doms = c5_domains(48, seed=5)
    _, pb = channel()
    init, steps = (P.INIT_PERTURBED, 31, 1e-4), 1000
    b = P.Batch(19, "f32s", doms, pb)
    ref = P.Batch(19, "f64", doms, pb)
    for x in (b, ref):
        x.initialize(1.0, (0, 0, 0), init)
        x.step(steps, steps_per_graph=20)
    worst = 0.0
    for d in range(len(doms)):
        _, u = b.moments(d)
        _, ur = ref.moments(d)
        fl = (slice(1, -1),)
        e = float(np.sqrt(((u[fl] - ur[fl]) ** 2).sum() / (ur[fl] ** 2).sum()))
        worst = max(worst, e)
        assert e <= 2.0 ** -24 * steps, (d, doms[d], e)
    print(f"largest per-domain velocity L2 {worst:.3e} (bound {2.0 ** -24 * steps:.3e})", flush=True)
    b.close()
    ref.close()
