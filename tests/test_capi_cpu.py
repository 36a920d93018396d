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


@pytest.mark.parametrize("mode", [capi.LITERAL, capi.FUSED])
@pytest.mark.parametrize("dims", [(1 << 22, 1 << 22, 1 << 22),        # X*Y*Z wraps to 0 in 64 bits
                                  ((1 << 31) - 1,) * 3,
                                  (1 << 16, 1 << 16, 1 << 16)])        # 2^48 cells
def test_oversized_domains_are_refused_before_any_device_call(mode, dims):
    """Cell counts that no device holds are refused by their size alone,
    before a product of the edges can wrap and size a buffer too small for
    the kernels' indices.  The check needs no GPU."""
    with pytest.raises(P.InvalidArgument, match="too large"):
        P.State(19, *dims, variant=capi.F64, mode=mode)


def test_wrapper_refuses_edges_ctypes_would_truncate():
    """2^32 + 4 would reach the int32 descriptor as 4."""
    with pytest.raises(P.InvalidArgument, match="too large"):
        P.State(19, (1 << 32) + 4, 4, 4, variant=capi.F64, mode=capi.FUSED)
    dom = {"x": 4, "y": (1 << 32) + 4, "z": 4, "tau": 0.8, "force": (0, 0, 0)}
    with pytest.raises(P.InvalidArgument, match="too large"):
        capi.Batch(19, capi.F32S, [dom], capi.channel())


def test_literal_row_length_limit():
    """Literal states keep the halo'd row length X + 2 in an int."""
    with pytest.raises(P.InvalidArgument, match="too large"):
        P.State(19, (1 << 31) - 2, 1, 1, variant=capi.F64, mode=capi.LITERAL)


def test_fused_cell_limit_is_2_to_the_32():
    """Fused states number cells in 32 bits: 2^32 cells are refused by size."""
    with pytest.raises(P.InvalidArgument, match="too large"):
        P.State(19, 1 << 16, 1 << 8, 1 << 8, variant=capi.F32S, mode=capi.FUSED)


def test_oversized_batch_domain_is_refused():
    """A batch domain whose edges multiply past 2^64 (to 0) is refused, not
    taken for an empty domain."""
    dom = {"x": 1 << 22, "y": 1 << 22, "z": 1 << 22, "tau": 0.8, "force": (0, 0, 0)}
    with pytest.raises(P.InvalidArgument, match="too large"):
        capi.Batch(19, capi.F32S, [dom], capi.channel())


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
