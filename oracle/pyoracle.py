"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by __graft_entry__.smoke() (to compare with) and by
bench.py's cpu_baseline leg.  The product (plbpd_amd/) never imports it.

Arrays follow the reference's AoS order: literal mode uses the halo'd field
layout of Field.hpp:91-93 ((Z+2, Y+2, X+2, vlen)), corrected mode the same
order without the halo ((Z, Y, X, vlen)).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(_HERE, "_build", "liborc.so")
BASE_SO = os.path.join(_HERE, "_build", "libcpubase.so")
REF_SO = os.path.join(_HERE, "_ref", "libplbpd_ref.so")

EAST, WEST, NORTH, SOUTH, UP, DOWN = range(6)
UNSET, PERIODIC, NOSLIP = 0, 1, 2
VARIANTS = {"f64": 0, "f32": 1, "f64a32": 2, "f32s": 3, "f16s": 4}
STORAGE_DTYPE = {0: np.float64, 1: np.float32, 2: np.float64, 3: np.float32, 4: np.uint16}


class OrcBC(C.Structure):
    _fields_ = [("type", C.c_int * 6), ("has_vel", C.c_int * 6),
                ("vel", (C.c_double * 3) * 6), ("rho_s", C.c_double)]


class OrcInit(C.Structure):
    _fields_ = [("kind", C.c_int), ("seed", C.c_ulonglong), ("amp", C.c_double)]


@dataclass
class Boundary:
    """Per-face types and velocities in reference face order."""
    type: list = field(default_factory=lambda: [UNSET] * 6)
    vel: dict = field(default_factory=dict)  # face -> (ux, uy, uz)
    rho_s: float = 1.0

    def to_c(self) -> OrcBC:
        b = OrcBC()
        for f in range(6):
            b.type[f] = self.type[f]
            b.has_vel[f] = 1 if f in self.vel else 0
            v = self.vel.get(f, (0.0, 0.0, 0.0))
            for i in range(3):
                b.vel[f][i] = float(v[i])
        b.rho_s = float(self.rho_s)
        return b


def channel_bc(walls_vel_up=(0.0, 0.0, 0.0), periodic_y=True, rho_s=1.0) -> Boundary:
    """Periodic east/west (and north/south), no-slip up/down."""
    t = [PERIODIC, PERIODIC, PERIODIC if periodic_y else UNSET,
         PERIODIC if periodic_y else UNSET, NOSLIP, NOSLIP]
    return Boundary(type=t, vel={UP: tuple(walls_vel_up), DOWN: (0.0, 0.0, 0.0)}, rho_s=rho_s)


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _dptr(a):
    return None if a is None else (C.c_double * 3)(*[float(v) for v in a])


class Oracle:
    def __init__(self, path: str = ORC_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle not built: {path} (run make -C oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.orc_tables.argtypes = [i, vp, vp, vp, vp]
        L.orc_feq_f64.argtypes = [i, d, vp, vp]
        L.orc_feq_f32.argtypes = [i, C.c_float, vp, vp]
        L.orc_sub2ind.argtypes = [C.c_uint] * 9
        L.orc_sub2ind.restype = C.c_uint
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_unit_uniform.argtypes = [C.c_ulonglong, C.c_ulonglong, i]
        L.orc_unit_uniform.restype = d
        L.orc_set_num_threads.argtypes = [i]
        L.orc_num_threads.restype = i
        for suf in ("f64", "f32"):
            getattr(L, f"orc_lit_init_{suf}").argtypes = [i, i, i, i, d, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, f"orc_lit_collide_{suf}").argtypes = [i, i, i, i, d, vp, vp, vp, vp]
            getattr(L, f"orc_lit_push_{suf}").argtypes = [i, i, i, i, vp, vp]
            getattr(L, f"orc_lit_noslip_{suf}").argtypes = [i, i, i, i, vp, vp, vp]
            getattr(L, f"orc_lit_periodic_{suf}").argtypes = [i, i, i, i, vp, vp]
            getattr(L, f"orc_lit_moments_{suf}").argtypes = [i, i, i, i, vp, vp, vp]
            getattr(L, f"orc_lit_reset_{suf}").argtypes = [i, i, i, i, vp, vp, vp]
            getattr(L, f"orc_lit_run_{suf}").argtypes = [i, i, i, i, d, vp, vp, i, vp, vp, vp, vp]
        L.orc_storage_bytes.argtypes = [i]
        L.orc_storage_bytes.restype = C.c_size_t
        L.orc_cor_flags.argtypes = [i, i, i, i, vp, vp, vp]
        L.orc_cor_init.argtypes = [i, i, i, i, i, d, vp, vp, vp]
        L.orc_cor_run.argtypes = [i, i, i, i, i, d, vp, vp, vp, i, vp, vp]
        L.orc_cor_moments.argtypes = [i, i, i, i, i, vp, vp, vp, vp, vp]
        L.orc_cor_populations.argtypes = [i, i, i, i, i, vp, vp]

    # ---- tables / feq ----
    def tables(self, q: int):
        c = np.zeros((q, 3), np.int32)
        w = np.zeros(q, np.float64)
        wf = np.zeros(q, np.float32)
        rev = np.zeros(q, np.int32)
        rc = self.lib.orc_tables(q, _ptr(c), _ptr(w), _ptr(wf), _ptr(rev))
        if rc:
            raise ValueError(f"unknown lattice q={q}")
        return c, w, wf, rev

    def feq(self, q, rho, u, dtype=np.float64):
        out = np.zeros(q, dtype)
        if dtype == np.float64:
            uu = np.asarray(u, np.float64)
            self.lib.orc_feq_f64(q, float(rho), _ptr(uu), _ptr(out))
        else:
            uu = np.asarray(u, np.float32)
            self.lib.orc_feq_f32(q, C.c_float(rho), _ptr(uu), _ptr(out))
        return out

    def sub2ind(self, zlen, ylen, xlen, vlen, nbuf, z, y, x, v):
        return self.lib.orc_sub2ind(zlen, ylen, xlen, vlen, nbuf, z, y, x, v)

    def f32_to_f16(self, f):
        return self.lib.orc_f32_to_f16(C.c_float(f))

    def f16_to_f32(self, h):
        return self.lib.orc_f16_to_f32(int(h))

    def unit_uniform(self, seed, cell, comp):
        return self.lib.orc_unit_uniform(seed, cell, comp)

    def set_threads(self, n):
        self.lib.orc_set_num_threads(int(n))

    # ---- literal mode ----
    def lit_state(self, q, X, Y, Z, dtype=np.float64):
        shp = (Z + 2, Y + 2, X + 2)
        return {"rho": np.zeros(shp, dtype), "u": np.zeros(shp + (3,), dtype),
                "n": np.zeros(shp + (q,), dtype), "ntmp": np.zeros(shp + (q,), dtype),
                "q": q, "X": X, "Y": Y, "Z": Z, "suf": "f64" if dtype == np.float64 else "f32"}

    def lit_init(self, st, rho0=1.0, u0=(0, 0, 0), init=None, bc: Boundary | None = None):
        ini = OrcInit(*(init or (0, 0, 0.0)))
        u = np.asarray(u0, np.float64)
        fn = getattr(self.lib, f"orc_lit_init_{st['suf']}")
        rc = fn(st["q"], st["X"], st["Y"], st["Z"], float(rho0), _ptr(u), C.byref(ini),
                C.byref(bc.to_c()) if bc else None, _ptr(st["rho"]), _ptr(st["u"]),
                _ptr(st["n"]), _ptr(st["ntmp"]))
        assert rc == 0

    def _dims(self, st):
        return st["q"], st["X"], st["Y"], st["Z"]

    def lit_collide(self, st, tau, F):
        fn = getattr(self.lib, f"orc_lit_collide_{st['suf']}")
        assert fn(*self._dims(st), float(tau), _dptr(F), _ptr(st["rho"]), _ptr(st["u"]), _ptr(st["n"])) == 0

    def lit_push(self, st):
        """push n -> ntmp, then swap (Streaming.cpp:93)."""
        fn = getattr(self.lib, f"orc_lit_push_{st['suf']}")
        assert fn(*self._dims(st), _ptr(st["n"]), _ptr(st["ntmp"])) == 0
        st["n"], st["ntmp"] = st["ntmp"], st["n"]

    def lit_noslip(self, st, bc: Boundary):
        fn = getattr(self.lib, f"orc_lit_noslip_{st['suf']}")
        return fn(*self._dims(st), C.byref(bc.to_c()), _ptr(st["rho"]), _ptr(st["n"]))

    def lit_periodic(self, st, bc: Boundary):
        fn = getattr(self.lib, f"orc_lit_periodic_{st['suf']}")
        assert fn(*self._dims(st), C.byref(bc.to_c()), _ptr(st["n"])) == 0

    def lit_moments(self, st):
        fn = getattr(self.lib, f"orc_lit_moments_{st['suf']}")
        assert fn(*self._dims(st), _ptr(st["n"]), _ptr(st["rho"]), _ptr(st["u"])) == 0

    def lit_reset(self, st, bc: Boundary):
        fn = getattr(self.lib, f"orc_lit_reset_{st['suf']}")
        assert fn(*self._dims(st), C.byref(bc.to_c()), _ptr(st["rho"]), _ptr(st["u"])) == 0

    def lit_run(self, st, tau, F, bc: Boundary, nsteps):
        fn = getattr(self.lib, f"orc_lit_run_{st['suf']}")
        rc = fn(*self._dims(st), float(tau), _dptr(F), C.byref(bc.to_c()), int(nsteps),
                _ptr(st["rho"]), _ptr(st["u"]), _ptr(st["n"]), _ptr(st["ntmp"]))
        if rc < 0:
            raise RuntimeError(f"orc_lit_run failed: {rc}")
        if rc == 1:
            st["n"], st["ntmp"] = st["ntmp"], st["n"]

    # ---- corrected mode ----
    def cor_flags(self, q, X, Y, Z, bc: Boundary, solid=None):
        fl = np.zeros((Z, Y, X), np.uint8)
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        rc = self.lib.orc_cor_flags(q, X, Y, Z, C.byref(bc.to_c()), None if s is None else _ptr(s), _ptr(fl))
        if rc:
            raise ValueError(f"invalid corrected-mode boundary ({rc})")
        return fl

    def cor_state(self, variant, q, X, Y, Z):
        v = VARIANTS[variant] if isinstance(variant, str) else variant
        dt = STORAGE_DTYPE[v]
        return {"v": v, "q": q, "X": X, "Y": Y, "Z": Z,
                "f": np.zeros((Z, Y, X, q), dt), "g": np.zeros((Z, Y, X, q), dt)}

    def cor_init(self, st, rho0=1.0, u0=(0, 0, 0), init=None):
        ini = OrcInit(*(init or (0, 0, 0.0)))
        u = np.asarray(u0, np.float64)
        assert self.lib.orc_cor_init(st["v"], st["q"], st["X"], st["Y"], st["Z"], float(rho0),
                                     _ptr(u), C.byref(ini), _ptr(st["f"])) == 0

    def cor_run(self, st, tau, F, bc: Boundary, nsteps, solid=None):
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        rc = self.lib.orc_cor_run(st["v"], st["q"], st["X"], st["Y"], st["Z"], float(tau), _dptr(F),
                                  C.byref(bc.to_c()), None if s is None else _ptr(s), int(nsteps),
                                  _ptr(st["f"]), _ptr(st["g"]))
        if rc < 0:
            raise RuntimeError(f"orc_cor_run failed: {rc}")
        if rc == 1:
            st["f"], st["g"] = st["g"], st["f"]

    def cor_moments(self, st, bc: Boundary, solid=None):
        Z, Y, X = st["Z"], st["Y"], st["X"]
        rho = np.zeros((Z, Y, X), np.float64)
        u = np.zeros((Z, Y, X, 3), np.float64)
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        rc = self.lib.orc_cor_moments(st["v"], st["q"], X, Y, Z, C.byref(bc.to_c()),
                                      None if s is None else _ptr(s), _ptr(st["f"]), _ptr(rho), _ptr(u))
        assert rc == 0
        return rho, u

    def cor_populations(self, st):
        out = np.zeros(st["f"].shape, np.float64)
        assert self.lib.orc_cor_populations(st["v"], st["q"], st["X"], st["Y"], st["Z"],
                                            _ptr(st["f"]), _ptr(out)) == 0
        return out


class RefProbe:
    """The reference's own LBModel / EqlbDist / Field / Domain code (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        vp = C.c_void_p
        self.lib.ref_tables.argtypes = [C.c_char_p, vp, vp, vp, vp]
        self.lib.ref_feq.argtypes = [C.c_char_p, C.c_float, vp, vp]
        self.lib.ref_field.argtypes = [C.c_int] + [C.c_uint32] * 8 + [vp, vp]
        self.lib.ref_domain.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                        C.c_float, vp, vp, vp, vp, vp]

    def tables(self, name: str):
        c = np.zeros((27, 3), np.int32)
        w = np.zeros(27, np.float32)
        rev = np.zeros(27, np.uint32)
        cs2 = np.zeros(1, np.float32)
        q = self.lib.ref_tables(name.encode(), _ptr(c), _ptr(w), _ptr(rev), _ptr(cs2))
        if q < 0:
            raise ValueError(name)
        return c[:q], w[:q], rev[:q], float(cs2[0])

    def feq(self, name, rho, u):
        out = np.zeros(27, np.float32)
        uu = np.asarray(u, np.float32)
        q = self.lib.ref_feq(name.encode(), C.c_float(rho), _ptr(uu), _ptr(out))
        return out[:q]

    def field(self, nbuf, z, y, x, v, qz, qy, qx, qv):
        idx = np.zeros(1, np.uint32)
        ext = np.zeros(6, np.uint32)
        rc = self.lib.ref_field(nbuf, z, y, x, v, qz, qy, qx, qv, _ptr(idx), _ptr(ext))
        return rc, int(idx[0]), ext

    def domain(self, name, x, y, z, nu, rho, u0, F):
        dims = np.zeros(3, np.uint64)
        nu_o = np.zeros(1, np.float32)
        rho_o = np.zeros(1, np.float32)
        u = np.asarray(u0, np.float32)
        f = np.asarray(F, np.float32)
        rc = self.lib.ref_domain(name.encode(), x, y, z, C.c_float(nu), C.c_float(rho), _ptr(u), _ptr(f),
                                 _ptr(dims), _ptr(nu_o), _ptr(rho_o))
        return rc, tuple(int(v) for v in dims), float(nu_o[0]), float(rho_o[0])


class CpuBaseline:
    """The reference's FP32 staged CPU step restated with OpenMP + AVX2."""

    def __init__(self, path: str = BASE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.lib.cpubase_run.argtypes = [C.c_int] * 4 + [C.c_float, C.c_void_p, C.c_int, C.c_int,
                                                        C.c_int, C.c_void_p]
        self.lib.cpubase_run.restype = C.c_double
        self.lib.cpubase_max_threads.restype = C.c_int

    def run(self, q, X, Y, Z, nu, F, warmup, steps, threads):
        phase = np.zeros(6, np.float64)
        f = np.asarray(F, np.float32)
        t = self.lib.cpubase_run(q, X, Y, Z, C.c_float(nu), _ptr(f), warmup, steps, threads, _ptr(phase))
        if t < 0:
            raise ValueError("bad cpu baseline arguments")
        return t, phase

    def max_threads(self):
        return self.lib.cpubase_max_threads()
