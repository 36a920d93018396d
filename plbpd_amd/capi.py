"""ctypes binding of the plbpd HIP C ABI (include/plbpd_hip.h).

The library is loaded from plbpd_amd/lib/libplbpd_hip.so (built in-tree by
``make -C plbpd_amd``).  There is no CPU fallback: if the library is missing
every entry point raises.  Status codes map to the exception types the
reference throws (see the header): invalid_argument -> InvalidArgument,
logic_error -> LogicError, out_of_range -> OutOfRange.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libplbpd_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "plbpd_hip.h")

F64, F32, F64_A32, F32S, F16S = 0, 1, 2, 3, 4
VARIANT_NAMES = {"f64": F64, "f32": F32, "f64a32": F64_A32, "f32s": F32S, "f16s": F16S}
STORAGE_BYTES = {F64: 8, F32: 4, F64_A32: 8, F32S: 4, F16S: 2}
STORAGE_DTYPE = {F64: np.float64, F32: np.float32, F64_A32: np.float64, F32S: np.float32, F16S: np.uint16}
LITERAL, FUSED, FUSED_AA = 0, 1, 2
EAST, WEST, NORTH, SOUTH, UP, DOWN = range(6)
UNSET, PERIODIC, NOSLIP = 0, 1, 2
FIELD_RHO, FIELD_U, FIELD_N, FIELD_NTMP = range(4)
LAYOUT_REFERENCE, LAYOUT_COMPACT = 0, 1
DTYPE_F64, DTYPE_F32 = 0, 1
INIT_UNIFORM, INIT_PERTURBED, INIT_TAYLOR_GREEN = 0, 1, 2
(TUNE_BLOCK_X, TUNE_NT_LOAD, TUNE_NT_STORE, TUNE_SCHEME, TUNE_TILE_ORDER, TUNE_VECTOR, TUNE_XCD, TUNE_INDEX64,
 TUNE_LITERAL_FUSE, TUNE_TWO_STEP, TUNE_TWO_ZCHUNK, TUNE_TWO_TILE_X) = range(12)
SCHEME_AUTO, SCHEME_PUSH, SCHEME_PULL, SCHEME_AA, SCHEME_PULL_TWO = 0, 1, 2, 3, 4
STAGES = ["collide", "stream", "noslip", "periodicity", "moments", "velocity_reset",
          "density_reset", "fused"]


class PlbError(RuntimeError):
    code = -1


class InvalidArgument(PlbError, ValueError):
    code = 1


class LogicError(PlbError):
    code = 2


class OutOfRange(PlbError, IndexError):
    code = 3


class HipError(PlbError):
    code = 4


class NoDevice(PlbError):
    code = 5


class OutOfMemory(PlbError, MemoryError):
    code = 6


_ERRORS = {c.code: c for c in (InvalidArgument, LogicError, OutOfRange, HipError, NoDevice, OutOfMemory)}


class StateDesc(C.Structure):
    _fields_ = [("q", C.c_int32), ("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32),
                ("variant", C.c_int32), ("mode", C.c_int32), ("device", C.c_int32),
                ("pop_pad", C.c_int32)]


class BoundaryC(C.Structure):
    _fields_ = [("type", C.c_int32 * 6), ("has_vel", C.c_int32 * 6),
                ("vel", (C.c_double * 3) * 6), ("rho_s", C.c_double)]


class InitSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("amp", C.c_double)]


class DomainDesc(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("tau", C.c_double),
                ("force", C.c_double * 3)]


def boundary(types, vel=None, rho_s=1.0) -> BoundaryC:
    """types: 6 face types (east, west, north, south, up, down); vel: {face: (ux,uy,uz)}."""
    b = BoundaryC()
    vel = vel or {}
    for f in range(6):
        b.type[f] = int(types[f])
        b.has_vel[f] = 1 if f in vel else 0
        v = vel.get(f, (0.0, 0.0, 0.0))
        for i in range(3):
            b.vel[f][i] = float(v[i])
    b.rho_s = float(rho_s)
    return b


def channel(up_velocity=(0.0, 0.0, 0.0), periodic_y=True, rho_s=1.0) -> BoundaryC:
    """Walls on up/down (no-slip), periodic east/west and north/south."""
    t = [PERIODIC, PERIODIC, PERIODIC if periodic_y else UNSET, PERIODIC if periodic_y else UNSET,
         NOSLIP, NOSLIP]
    return boundary(t, {UP: tuple(up_velocity), DOWN: (0.0, 0.0, 0.0)}, rho_s)


_lib = None


def exported_names():
    """Every function the public header declares."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(plb_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"plbpd HIP library not built: {LIB_PATH} (run make -C plbpd_amd)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    L.plb_last_error.restype = C.c_char_p
    L.plb_device_info.argtypes = [i, C.c_char_p, sz, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
    L.plb_lattice_tables.argtypes = [i, vp, vp, vp]
    L.plb_state_create.argtypes = [C.POINTER(StateDesc), C.POINTER(vp)]
    L.plb_state_destroy.argtypes = [vp]
    L.plb_state_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.plb_upload.argtypes = [vp, i, i, i, vp]
    L.plb_download.argtypes = [vp, i, i, i, vp]
    L.plb_download_raw.argtypes = [vp, vp, sz]
    L.plb_domain_initialize.argtypes = [vp, d, vp, C.POINTER(InitSpec), C.POINTER(BoundaryC)]
    L.plb_set_boundary.argtypes = [vp, C.POINTER(BoundaryC), vp]
    L.plb_collide_srt.argtypes = [vp, d, vp]
    L.plb_stream.argtypes = [vp, C.c_char_p]
    for n in ("plb_apply_noslip", "plb_apply_noslip_noalloc", "plb_apply_periodicity", "plb_reset",
              "plb_reset_velocity", "plb_reset_density"):
        getattr(L, n).argtypes = [vp, C.POINTER(BoundaryC)]
    L.plb_calc_moments.argtypes = [vp]
    L.plb_configure.argtypes = [vp, d, vp, C.POINTER(BoundaryC)]
    L.plb_step.argtypes = [vp, i]
    L.plb_step_graph.argtypes = [vp, i, i]
    L.plb_global_moments.argtypes = [vp, vp]
    L.plb_synchronize.argtypes = [vp]
    L.plb_enable_timers.argtypes = [vp, i]
    L.plb_stage_times.argtypes = [vp, vp]
    L.plb_time_steps.argtypes = [vp, i, i, C.POINTER(C.c_double)]
    L.plb_bench_copy.argtypes = [i, sz, i, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.plb_bench_pattern.argtypes = [i] * 9 + [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.plb_set_tuning.argtypes = [vp, i, i]
    L.plb_step_scheme.argtypes = [vp, C.POINTER(C.c_int)]
    if hasattr(L, "plb_batch_create"):
        L.plb_batch_create.argtypes = [i, i, i, C.POINTER(DomainDesc), C.POINTER(BoundaryC), i, C.POINTER(vp)]
        L.plb_batch_destroy.argtypes = [vp]
        L.plb_batch_initialize.argtypes = [vp, d, vp, C.POINTER(InitSpec)]
        L.plb_batch_set_tiling.argtypes = [vp, i, i]
        L.plb_batch_get_tiling.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.plb_batch_autotune.argtypes = [vp, i, vp, i, C.POINTER(C.c_int)]
        L.plb_batch_set_mapping.argtypes = [vp, i]
        L.plb_batch_set_scheme.argtypes = [vp, i]
        L.plb_batch_get_mapping.argtypes = [vp, C.POINTER(C.c_int)]
        L.plb_batch_step.argtypes = [vp, i, i]
        L.plb_batch_step_per_domain.argtypes = [vp, i]
        L.plb_batch_download_moments.argtypes = [vp, i, vp, vp]
        L.plb_batch_cells.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.plb_batch_synchronize.argtypes = [vp]
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        msg = lib().plb_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, PlbError)(msg)


def _p(a: np.ndarray):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def _vec3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def device_info(device=0):
    name = C.create_string_buffer(256)
    ncu = C.c_int(0)
    mem = C.c_size_t(0)
    check(lib().plb_device_info(device, name, 256, C.byref(ncu), C.byref(mem)))
    return {"name": name.value.decode(), "num_cu": ncu.value, "hbm_bytes": mem.value}


def bench_copy(nbytes=1 << 31, iters=20, device=0):
    """Device copy bandwidth (GB/s, read + write) and ms per copy."""
    g, ms = C.c_double(), C.c_double()
    check(lib().plb_bench_copy(device, int(nbytes), int(iters), C.byref(g), C.byref(ms)))
    return g.value, ms.value


PATTERN_COPY, PATTERN_PULL, PATTERN_PULL_SOLIDS = 0, 1, 2


def bench_pattern(q, X, Y, Z, sbytes, vec, pattern, iters=20, device=0):
    """Memory-pattern ceiling of the fused step (plb_bench_pattern): the
    step's algorithmic GB/s at the pattern's mean launch time, and ms."""
    g, ms = C.c_double(), C.c_double()
    check(lib().plb_bench_pattern(device, q, X, Y, Z, sbytes, vec, pattern, int(iters), C.byref(g), C.byref(ms)))
    return g.value, ms.value


def lattice_tables(q: int):
    c = np.zeros((q, 3), np.int32)
    w = np.zeros(q, np.float64)
    rev = np.zeros(q, np.uint32)
    check(lib().plb_lattice_tables(q, _p(c), _p(w), _p(rev)))
    return c, w, rev


def _check_edges(*edges):
    """The descriptors hold int32 edges, and ctypes would truncate a wider
    one silently; the library refuses zero and oversized extents itself."""
    for e in edges:
        if int(e) > 0x7FFFFFFF:
            raise InvalidArgument("domain too large")


class State:
    """A device lattice (the device side of SimData) with one method per
    reference operator call."""

    def __init__(self, q, X, Y, Z, variant=F64, mode=LITERAL, device=0, pop_pad=-1):
        if isinstance(variant, str):
            variant = VARIANT_NAMES[variant]
        _check_edges(X, Y, Z)
        self.q, self.X, self.Y, self.Z = q, X, Y, Z
        self.variant, self.mode = variant, mode
        d = StateDesc(q, X, Y, Z, variant, mode, device, pop_pad)
        h = C.c_void_p()
        check(lib().plb_state_create(C.byref(d), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().plb_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- shapes ---
    def shape(self, layout=LAYOUT_REFERENCE):
        if layout == LAYOUT_REFERENCE:
            return (self.Z + 2, self.Y + 2, self.X + 2)
        return (self.Z, self.Y, self.X)

    def nbytes(self):
        b = C.c_size_t()
        check(lib().plb_state_bytes(self.h, C.byref(b)))
        return b.value

    # --- data ---
    def _field_shape(self, field, layout):
        vlen = {FIELD_RHO: 1, FIELD_U: 3, FIELD_N: self.q, FIELD_NTMP: self.q}[field]
        return self.shape(layout) + ((vlen,) if vlen > 1 else ())

    def upload(self, field, arr, layout=LAYOUT_REFERENCE):
        a = np.ascontiguousarray(arr)
        if a.dtype not in (np.float64, np.float32):
            raise TypeError("float32 or float64 arrays only")
        shp = self._field_shape(field, layout)
        if a.size != int(np.prod(shp)):
            raise ValueError(f"array of {a.size} values for a field of shape {shp}")
        dt = DTYPE_F64 if a.dtype == np.float64 else DTYPE_F32
        check(lib().plb_upload(self.h, field, layout, dt, _p(a)))

    def download(self, field, layout=LAYOUT_REFERENCE, dtype=np.float64, out=None):
        shp = self._field_shape(field, layout)
        if out is None:
            out = np.zeros(shp, dtype)
        if out.dtype not in (np.float64, np.float32) or out.size != int(np.prod(shp)):
            raise ValueError(f"out must be a float32 or float64 array of shape {shp}")
        dt = DTYPE_F64 if out.dtype == np.float64 else DTYPE_F32
        check(lib().plb_download(self.h, field, layout, dt, _p(out)))
        return out

    def download_raw(self):
        out = np.zeros((self.Z, self.Y, self.X, self.q), STORAGE_DTYPE[self.variant])
        check(lib().plb_download_raw(self.h, _p(out), out.nbytes))
        return out

    # --- set-up ---
    def initialize(self, rho0=1.0, u0=(0, 0, 0), init=None, bc: BoundaryC | None = None):
        spec = InitSpec(*(init or (INIT_UNIFORM, 0, 0.0)))
        check(lib().plb_domain_initialize(self.h, float(rho0), _vec3(u0), C.byref(spec),
                                          C.byref(bc) if bc is not None else None))

    def set_boundary(self, bc: BoundaryC, mask=None):
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if m is not None and m.size != self.X * self.Y * self.Z:
            raise ValueError(f"mask of {m.size} bytes for {self.Z} x {self.Y} x {self.X} cells")
        check(lib().plb_set_boundary(self.h, C.byref(bc), None if m is None else _p(m)))

    # --- reference operator calls (PLBPD.cpp:24-29) ---
    def collide(self, tau, force):
        check(lib().plb_collide_srt(self.h, float(tau), _vec3(force)))

    def stream(self, stream_type="push"):
        check(lib().plb_stream(self.h, stream_type.encode()))

    def apply_noslip(self, bc):
        check(lib().plb_apply_noslip(self.h, C.byref(bc)))

    def apply_periodicity(self, bc):
        check(lib().plb_apply_periodicity(self.h, C.byref(bc)))

    def calc_moments(self):
        check(lib().plb_calc_moments(self.h))

    def reset(self, bc):
        check(lib().plb_reset(self.h, C.byref(bc)))

    # --- whole steps ---
    def configure(self, tau, force, bc):
        check(lib().plb_configure(self.h, float(tau), _vec3(force), C.byref(bc)))

    def step(self, n=1):
        check(lib().plb_step(self.h, int(n)))

    def step_graph(self, n, steps_per_graph=20):
        check(lib().plb_step_graph(self.h, int(n), int(steps_per_graph)))

    def time_steps(self, n, steps_per_graph=0):
        """Milliseconds for n steps, timed with HIP events on the state's stream."""
        ms = C.c_double()
        check(lib().plb_time_steps(self.h, int(n), int(steps_per_graph), C.byref(ms)))
        return ms.value

    def set_tuning(self, key, value):
        check(lib().plb_set_tuning(self.h, int(key), int(value)))

    def scheme(self):
        """'push', 'pull', 'pull2' (pull, two steps per launch) or 'aa' (one
        array): the streaming scheme of the next fused steps."""
        v = C.c_int()
        check(lib().plb_step_scheme(self.h, C.byref(v)))
        return {SCHEME_PULL: "pull", SCHEME_AA: "aa", SCHEME_PULL_TWO: "pull2"}.get(v.value, "push")

    def steps_per_launch(self):
        """Steps one launch of the fused kernel advances (2 for 'pull2')."""
        return 2 if self.scheme() == "pull2" else 1

    def global_moments(self):
        out = np.zeros(4, np.float64)
        check(lib().plb_global_moments(self.h, _p(out)))
        return out

    def synchronize(self):
        check(lib().plb_synchronize(self.h))

    def enable_timers(self, on=True):
        check(lib().plb_enable_timers(self.h, 1 if on else 0))

    def stage_times(self):
        out = np.zeros(8, np.float64)
        check(lib().plb_stage_times(self.h, _p(out)))
        return dict(zip(STAGES, out.tolist()))

    # --- convenience ---
    def moments(self, layout=LAYOUT_COMPACT):
        return self.download(FIELD_RHO, layout), self.download(FIELD_U, layout)


class Batch:
    """Many independent domains advanced together (one launch per step)."""

    def __init__(self, q, variant, domains, bc: BoundaryC, device=0):
        if isinstance(variant, str):
            variant = VARIANT_NAMES[variant]
        arr = (DomainDesc * len(domains))()
        for i, dm in enumerate(domains):
            _check_edges(dm["x"], dm["y"], dm["z"])
            arr[i].x, arr[i].y, arr[i].z = dm["x"], dm["y"], dm["z"]
            arr[i].tau = float(dm["tau"])
            for j in range(3):
                arr[i].force[j] = float(dm["force"][j])
        self.domains = [dict(d) for d in domains]
        self.q, self.variant = q, variant
        h = C.c_void_p()
        check(lib().plb_batch_create(q, variant, len(domains), arr, C.byref(bc), device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().plb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, rho0=1.0, u0=(0, 0, 0), init=None):
        spec = InitSpec(*(init or (INIT_UNIFORM, 0, 0.0)))
        check(lib().plb_batch_initialize(self.h, float(rho0), _vec3(u0), C.byref(spec)))

    def set_tiling(self, tile_cells, grid_blocks=0):
        check(lib().plb_batch_set_tiling(self.h, int(tile_cells), int(grid_blocks)))

    def tiling(self):
        a, b = C.c_int(), C.c_int()
        check(lib().plb_batch_get_tiling(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_scheme(self, scheme):
        check(lib().plb_batch_set_scheme(self.h, int(scheme)))

    def set_mapping(self, xcd_runs):
        check(lib().plb_batch_set_mapping(self.h, int(xcd_runs)))

    def mapping(self):
        m = C.c_int()
        check(lib().plb_batch_get_mapping(self.h, C.byref(m)))
        return m.value

    def autotune(self, steps=10, max_results=128):
        """Rows of (tile_cells, grid_blocks, mapping, ms_per_step)."""
        res = np.zeros((max_results, 4), np.float64)
        n = C.c_int(0)
        check(lib().plb_batch_autotune(self.h, int(steps), _p(res), max_results, C.byref(n)))
        return res[: n.value]

    def step(self, n, steps_per_graph=0):
        check(lib().plb_batch_step(self.h, int(n), int(steps_per_graph)))

    def step_per_domain(self, n):
        check(lib().plb_batch_step_per_domain(self.h, int(n)))

    def moments(self, d):
        dm = self.domains[d]
        rho = np.zeros((dm["z"], dm["y"], dm["x"]), np.float64)
        u = np.zeros((dm["z"], dm["y"], dm["x"], 3), np.float64)
        check(lib().plb_batch_download_moments(self.h, int(d), _p(rho), _p(u)))
        return rho, u

    def cells(self):
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().plb_batch_cells(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def synchronize(self):
        check(lib().plb_batch_synchronize(self.h))
