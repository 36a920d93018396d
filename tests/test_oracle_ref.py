"""The oracle against the reference's own code (oracle/_ref: LBModel.cpp,
EqlbDist.hpp, Field.hpp and Domain.cpp compiled from the reference tree).
Skipped where that library has not been built."""
import json
import os

import numpy as np
import pytest

from oracle.pyoracle import REF_SO, RefProbe

pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built")
G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module")
def ref():
    return RefProbe()


@pytest.mark.parametrize("name,q", [("D2Q9", 9), ("D3Q19", 19), ("D3Q27", 27)])
def test_tables_equal_reference_code(orc, ref, name, q):
    rc, rw, rrev, cs2 = ref.tables(name)
    c, w, wf, rev = orc.tables(q)
    assert np.array_equal(rc, c)
    assert np.array_equal(rw, wf)
    assert np.array_equal(rrev.astype(np.int32), rev)
    assert np.float32(cs2) == np.float32(1.0 / 3.0)


def test_unknown_model_throws(ref):
    with pytest.raises(ValueError):
        ref.tables("D3Q15")


@pytest.mark.parametrize("name,q", [("D2Q9", 9), ("D3Q19", 19), ("D3Q27", 27)])
def test_feq_f32_equals_reference_code(orc, ref, name, q):
    rng = np.random.default_rng(q)
    for _ in range(300):
        rho = np.float32(rng.uniform(0.5, 1.5))
        u = rng.uniform(-0.3, 0.3, 3).astype(np.float32)
        assert np.array_equal(orc.feq(q, rho, u, dtype=np.float32), ref.feq(name, float(rho), u))
    g = G["eqlbdist_d2q9"]
    assert np.array_equal(ref.feq("D2Q9", 1.0, g["u"]), np.array(g["expected"]).astype(np.float32))


def test_field_index_equals_reference_code(orc, ref):
    rng = np.random.default_rng(1)
    for _ in range(200):
        nbuf = int(rng.integers(0, 2))
        z, y, x = (int(v) for v in rng.integers(1, 9, 3))
        v = int(rng.integers(1, 28))
        qz, qy, qx = (int(rng.integers(0, d + 2 * nbuf)) for d in (z, y, x))
        qv = int(rng.integers(0, v))
        rc, idx, ext = ref.field(nbuf, z, y, x, v, qz, qy, qx, qv)
        assert rc == 0
        assert idx == orc.sub2ind(z, y, x, v, nbuf, qz, qy, qx, qv)
        assert ext.tolist() == [nbuf, z + nbuf, nbuf, y + nbuf, nbuf, x + nbuf]
    f = G["field"]["scalar_1buf"]
    for kv in f["known"]:
        rc, idx, ext = ref.field(1, f["zsize"], f["ysize"], f["xsize"], 1, *kv["zyx"], 0)
        assert idx == kv["value"] and ext.tolist() == f["extents"]
    assert ref.field(1, 0, 2, 2, 1, 0, 0, 0, 0)[0] == 1  # zero extent throws (Field.hpp:28-31)


def test_domain_getters(ref):
    g = G["domain"]
    rc, dims, nu, rho = ref.domain(g["model"], *g["dims_xyz"], g["nu"], g["rho"], g["vel"], g["force"])
    assert rc == 0 and dims == tuple(g["dims_xyz"]) and nu == g["nu"] and rho == g["rho"]
