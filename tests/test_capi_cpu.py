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


def _hostless_state(q=19, X=4, Y=3, Z=2):
    """A State object without a device state: enough for the host-side
    argument checks, which run before any library call."""
    from plbpd_amd import capi
    st = object.__new__(capi.State)
    st.q, st.X, st.Y, st.Z, st.h = q, X, Y, Z, None
    st.variant, st.mode = capi.F64, capi.FUSED
    return st


def test_wrapper_rejects_wrong_sized_arrays():
    """Arrays handed to the library as raw pointers are checked for size
    first, so a wrong shape raises instead of reading or writing past the
    host buffer."""
    import numpy as np
    from plbpd_amd import capi
    st = _hostless_state()
    with pytest.raises(ValueError):
        st.set_boundary(capi.channel(), np.zeros((2, 3, 3), np.uint8))
    with pytest.raises(ValueError):
        st.upload(capi.FIELD_RHO, np.zeros((4, 5, 5)))  # reference layout is (Z+2, Y+2, X+2)
    with pytest.raises(ValueError):
        st.download(capi.FIELD_U, capi.LAYOUT_COMPACT, out=np.zeros((2, 3, 4)))  # u has 3 components
    with pytest.raises(TypeError):
        st.upload(capi.FIELD_RHO, np.zeros((4, 5, 6), np.int32))
