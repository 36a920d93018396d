"""plbpd_amd: MI355X (gfx950) lattice-Boltzmann BGK step behind plbpd's operator interfaces.

The compute path is the HIP library plbpd_amd/lib/libplbpd_hip.so (C ABI in
include/plbpd_hip.h); the C++ drop-in classes with the reference's names live
in plbpd_amd/host/.  This module is the ctypes binding used by tests and
bench.py.
"""
from .capi import (  # noqa: F401
    F64, F32, F64_A32, F32S, F16S, LITERAL, FUSED, FUSED_AA,
    EAST, WEST, NORTH, SOUTH, UP, DOWN, UNSET, PERIODIC, NOSLIP,
    FIELD_RHO, FIELD_U, FIELD_N, FIELD_NTMP, LAYOUT_REFERENCE, LAYOUT_COMPACT,
    INIT_UNIFORM, INIT_PERTURBED, INIT_TAYLOR_GREEN, STORAGE_BYTES,
    State, Batch, boundary, channel, lattice_tables, device_info, lib,
    PlbError, InvalidArgument, LogicError, OutOfRange, HipError, NoDevice,
)
