"""The oracle against the vectors held by the reference's own unit tests
(tests/golden/reference_vectors.json), and its half-precision conversion
against numpy's IEEE float16."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def _w(text):
    a, b = text.rstrip(".").split("/")
    return float(a) / float(b.rstrip("."))


@pytest.mark.parametrize("name", ["D2Q9", "D3Q19", "D3Q27"])
def test_tables_match_testlbmodel(orc, name):
    g = G["lbmodel"][name]
    c, w, wf, rev = orc.tables(g["q"])
    assert c.tolist() == g["c"]
    assert rev.tolist() == g["rev"]
    exp = np.array([float(Fraction(t.replace(".", ""))) for t in g["w_text"]])
    assert np.array_equal(w, exp)  # the double literals 1./3. etc.
    assert np.array_equal(wf, exp.astype(np.float32))  # stored as float by the reference


def test_weights_and_moments_of_tables(orc):
    for q in (9, 19, 27):
        c, w, _, rev = orc.tables(q)
        assert abs(w.sum() - 1.0) < 1e-15
        assert np.allclose((w[:, None] * c).sum(0), 0.0, atol=1e-16)
        # second moment = cs2 delta_ij for the axes the lattice moves in
        m2 = np.einsum("k,ki,kj->ij", w, c, c)
        axes = [0, 2] if q == 9 else [0, 1, 2]
        for i in axes:
            assert abs(m2[i, i] - 1.0 / 3.0) < 1e-15
        assert np.array_equal(c[rev], -c)


def test_cs2inv_is_three():
    # Boundary.cpp:27: _cs2inv = 1.0f / float(1./3.) rounds to exactly 3.0f,
    # which the oracle's no-slip term uses (and 3.0 in double).
    cs2 = np.float32(1.0 / 3.0)
    assert np.float32(1.0) / cs2 == np.float32(3.0)
    assert 1.0 / (1.0 / 3.0) == 3.0


def test_feq_f32_matches_testeqlbdist(orc):
    g = G["eqlbdist_d2q9"]
    got = orc.feq(9, g["rho"], g["u"], dtype=np.float32)
    exp = np.array(g["expected"], np.float64).astype(np.float32)
    err = np.float32(0)
    for k in range(9):
        t = np.float32(exp[k] - got[k])
        err = np.float32(err + t * t)
    assert err == 0.0
    assert np.array_equal(got, exp)


def test_feq_f64_close_to_golden(orc):
    g = G["eqlbdist_d2q9"]
    got = orc.feq(9, g["rho"], g["u"], dtype=np.float64)
    assert np.allclose(got, np.array(g["expected"]), rtol=1e-7, atol=0)


def test_field_index_known_answers(orc):
    f = G["field"]["scalar_1buf"]
    for kv in f["known"]:
        z, y, x = kv["zyx"]
        assert orc.sub2ind(f["zsize"], f["ysize"], f["xsize"], 1, 1, z, y, x, 0) == kv["value"]
    # vector field with 0 halo layers: index grows by one in loop order
    v = G["field"]["vector_0buf"]
    ctr = 0
    for z in range(v["zsize"]):
        for y in range(v["ysize"]):
            for x in range(v["xsize"]):
                for k in range(v["vsize"]):
                    assert orc.sub2ind(v["zsize"], v["ysize"], v["xsize"], v["vsize"], 0, z, y, x, k) == ctr
                    ctr += 1


def test_f16_conversion_matches_numpy(orc):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(3e-3),
        rng.standard_normal(2000).astype(np.float32) * np.float32(1e3),
        (rng.random(2000).astype(np.float32) * np.float32(1e-6)),  # subnormal halves
        np.array([0.0, -0.0, 65504.0, 65519.996, 65520.0, 70000.0, -70000.0, 2.0 ** -24,
                  2.0 ** -25, 2.0 ** -25 * 1.0000001, 1.5 * 2.0 ** -24, 2.0 ** -14, 6.1e-5,
                  np.inf, -np.inf], np.float32),
    ])
    exp = vals.astype(np.float16).view(np.uint16)
    got = np.array([orc.f32_to_f16(float(v)) for v in vals], np.uint16)
    assert np.array_equal(got, exp)
    back = np.array([orc.f16_to_f32(int(h)) for h in exp], np.float32)
    assert np.array_equal(back, exp.view(np.float16).astype(np.float32))


def test_unit_uniform_range_and_spread(orc):
    v = np.array([orc.unit_uniform(42, c, k) for c in range(3000) for k in range(3)])
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.02 and abs(v.var() - 1 / 12) < 0.01
    assert len(set(v.tolist())) == len(v)
