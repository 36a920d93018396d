"""The C-ABI library on the CPU: it loads, exports every entry point the
public header declares, and its lattice tables equal the oracle's and the
reference's golden tables.  No call here touches a GPU."""
import ctypes
import json
import os

import numpy as np
import pytest

import plbpd_amd as P
from plbpd_amd import capi

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = capi.exported_names()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_loads():
    host = os.path.join(os.path.dirname(capi.LIB_PATH), "libplbpd_host.so")
    ctypes.CDLL(host)


@pytest.mark.parametrize("name,q", [("D2Q9", 9), ("D3Q19", 19), ("D3Q27", 27)])
def test_device_tables_equal_oracle_and_golden(orc, name, q):
    c, w, rev = P.lattice_tables(q)
    oc, ow, _, orev = orc.tables(q)
    assert np.array_equal(c, oc) and np.array_equal(w, ow) and np.array_equal(rev, orev)
    assert c.tolist() == G["lbmodel"][name]["c"] and rev.tolist() == G["lbmodel"][name]["rev"]


def test_unknown_lattice_is_logic_error():
    with pytest.raises(P.LogicError):
        P.lattice_tables(15)


def test_no_device_here_is_reported():
    try:
        P.device_info(0)
    except P.NoDevice:
        with pytest.raises(P.NoDevice):
            P.State(19, 4, 4, 4)
        return
    pytest.skip("a GPU is visible")
