"""ctypes wrapper of the FP64 CPU oracle (test infrastructure only).

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module, and only as the checker.  ``oracle/_ref`` holds the
reference's own dbyd2.f90+mtridag.f90 and dbyd1.f+tridag.f compiled with
amdflang (``make -C oracle ref``), used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF2 = None
_REF1 = None

DP = C.POINTER(C.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(DP)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_threads.restype = C.c_int
        L.orc_create.argtypes = [C.c_void_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_beta.restype = C.c_double
        L.orc_beta.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_background.argtypes = [C.c_void_p, C.c_int, DP, C.c_long]
        L.orc_set_profile.argtypes = [C.c_void_p, C.c_int, DP, C.c_int]
        L.orc_set_damping.argtypes = [C.c_void_p, DP, C.c_long]
        L.orc_set_filter.argtypes = [C.c_void_p, DP, DP]
        L.orc_set_source.argtypes = [C.c_void_p, C.c_int, DP, C.c_long, DP, C.c_int, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_get_state.argtypes = [C.c_void_p, C.c_int, C.c_int, DP]
        L.orc_put_state.argtypes = [C.c_void_p, C.c_int, C.c_int, DP]
        L.orc_norm.restype = C.c_double
        L.orc_norm.argtypes = [C.c_void_p, C.c_int]
        L.orc_dbyd2.argtypes = [DP, DP, C.c_long, C.c_int, C.c_int]
        L.orc_fft.argtypes = [DP, DP, C.c_int, C.c_int]
        L.orc_ddx.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_ddy.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_ddz.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, C.c_int, DP]
        L.orc_zfilter.argtypes = [DP, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_lowpass.argtypes = [DP, C.c_int, C.c_int, C.c_int, C.c_int, DP]
        L.orc_set_source_planes.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_long]
        L.orc_lagrange_eval.argtypes = [DP, C.c_long, C.c_int, C.c_int, C.c_double,
                                        C.c_double, C.c_int, DP]
        L.orc_compute_tt.restype = C.c_double
        L.orc_compute_tt.argtypes = [DP, DP, C.c_int, C.c_double]
        L.orc_inverse3.argtypes = [DP, DP]
        L.orc_snapshots_alloc.argtypes = [C.c_void_p, C.c_int]
        L.orc_snapshot_store.argtypes = [C.c_void_p, C.c_int]
        L.orc_snapshot_get.argtypes = [C.c_void_p, C.c_int, C.c_int, DP]
        L.orc_snapshot_put.argtypes = [C.c_void_p, C.c_int, C.c_int, DP]
        L.orc_kernel_begin.argtypes = [C.c_void_p, DP, C.c_int]
        L.orc_kernel_accumulate.argtypes = [C.c_void_p, C.c_int]
        L.orc_kernel_get.argtypes = [C.c_void_p, C.c_int, DP]
        L.orc_kernel_put.argtypes = [C.c_void_p, C.c_int, DP]
        L.orc_damp_field.argtypes = [C.c_void_p, DP, DP]
        L.orc_init_rhs.argtypes = ([C.c_int] * 6 + [C.c_double] * 2 + [DP] * 13)
        _LIB = L
    return _LIB


def ref_dbyd2():
    """The reference's DBYD2 (dbyd2.f90 + mtridag.f90), or None if not built."""
    global _REF2
    if _REF2 is None:
        path = os.path.join(_HERE, "_ref", "libref_dbyd2.so")
        if not os.path.exists(path):
            return None
        _REF2 = C.CDLL(path)
        _REF2.dbyd2_.argtypes = [DP, DP, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]
    return _REF2


def ref_dbyd1():
    global _REF1
    if _REF1 is None:
        path = os.path.join(_HERE, "_ref", "libref_dbyd1.so")
        if not os.path.exists(path):
            return None
        _REF1 = C.CDLL(path)
        _REF1.dbyd1_.argtypes = [DP, DP, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]
    return _REF1


def call_ref_dbyd2(A: np.ndarray, B: np.ndarray, ibc: int) -> np.ndarray:
    """A(N1,N2) <- DBYD2(B) in Fortran order; A holds caller boundary values."""
    R = ref_dbyd2()
    n1, n2 = B.shape[0], B.shape[1]
    A = np.asfortranarray(A.astype(np.float64))
    B = np.asfortranarray(B.astype(np.float64))
    R.dbyd2_(_dp(A), _dp(B), C.byref(C.c_int(n1)), C.byref(C.c_int(n2)), C.byref(C.c_int(ibc)))
    return A


def call_ref_dbyd1(A: np.ndarray, B: np.ndarray, ibc: int) -> np.ndarray:
    """A(N1,N2) <- DBYD1(B) along the first index, Fortran order; A holds
    caller boundary values (ibc 2, 3, 4)."""
    R = ref_dbyd1()
    n1, n2 = B.shape[0], B.shape[1]
    A = np.asfortranarray(A.astype(np.float64))
    B = np.asfortranarray(B.astype(np.float64))
    R.dbyd1_(_dp(A), _dp(B), C.byref(C.c_int(n1)), C.byref(C.c_int(n2)), C.byref(C.c_int(ibc)))
    return A


def dbyd2(A: np.ndarray, B: np.ndarray, ibc: int) -> np.ndarray:
    L = lib()
    n1, n2 = B.shape[0], B.shape[1]
    A = np.asfortranarray(A.astype(np.float64))
    B = np.asfortranarray(B.astype(np.float64))
    rc = L.orc_dbyd2(_dp(A), _dp(B), n1, n2, ibc)
    assert rc == 0
    return A


def ddx(f: np.ndarray, stretchx: float) -> np.ndarray:
    f = np.asfortranarray(f, dtype=np.float64)
    out = np.zeros_like(f, order="F")
    nx, ny, nz = f.shape
    lib().orc_ddx(_dp(f), _dp(out), nx, ny, nz, stretchx)
    return out


def ddy(f: np.ndarray, stretchy: float) -> np.ndarray:
    f = np.asfortranarray(f, dtype=np.float64)
    out = np.zeros_like(f, order="F")
    nx, ny, nz = f.shape
    lib().orc_ddy(_dp(f), _dp(out), nx, ny, nz, stretchy)
    return out


def ddz(f: np.ndarray, bc: int, stretch: np.ndarray, preset: np.ndarray | None = None):
    f = np.asfortranarray(f, dtype=np.float64)
    out = np.zeros_like(f, order="F") if preset is None else np.array(preset, dtype=np.float64, order="F")
    nx, ny, nz = f.shape
    st = np.ascontiguousarray(stretch, dtype=np.float64)
    lib().orc_ddz(_dp(f), _dp(out), nx, ny, nz, bc, _dp(st))
    return out


def zfilter(f: np.ndarray, kst: int) -> np.ndarray:
    f = np.array(f, dtype=np.float64, order="F")
    nx, ny, nz = f.shape
    lib().orc_zfilter(_dp(f), nx, ny, nz, kst)
    return f


def lowpass(f: np.ndarray, axis: int, decay: np.ndarray) -> np.ndarray:
    """Fermi low-pass along x (axis 0) or y (axis 1) with a decay table of
    n/2+1 entries (normx folded)."""
    f = np.array(f, dtype=np.float64, order="F")
    nx, ny, nz = f.shape
    d = np.ascontiguousarray(decay, dtype=np.float64)
    lib().orc_lowpass(_dp(f), nx, ny, nz, axis, _dp(d))
    return f


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    ri = np.ascontiguousarray(x.view(np.float64))
    out = np.zeros_like(ri)
    lib().orc_fft(_dp(ri), _dp(out), x.size, sign)
    return out.view(np.complex128)


def lagrange(vr: np.ndarray, time: int, timestep: float, cad: float, num_steps: int):
    """vr: (nxy, nt) samples, Fortran order (sample t at column t-1)."""
    vr = np.asfortranarray(vr, dtype=np.float64)
    nxy, nt = vr.shape
    out = np.zeros(nxy)
    lib().orc_lagrange_eval(_dp(vr), nxy, nt, time, timestep, cad, num_steps, _dp(out))
    return out


def compute_tt(u0: np.ndarray, u: np.ndarray, dt: float) -> float:
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_compute_tt(_dp(u0), _dp(u), u0.size, dt)


def inverse3(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    c = np.zeros((3, 3))
    lib().orc_inverse3(_dp(a), _dp(c))
    return c


def init_rhs(nx, ny, nz, is2d, npmlbot, npmltop, diml, dimc, z, c2, rho0, p0, g):
    """Derived background of mp_Initialize_RHS[_2d]; returns a dict."""
    N = nx * ny * nz
    f = lambda a: np.array(np.broadcast_to(a, (nx, ny, nz)), dtype=np.float64, order="F")
    z = np.ascontiguousarray(z, dtype=np.float64)
    c2, rho0, p0 = f(c2), f(rho0), f(p0)
    g = np.array(g, dtype=np.float64)
    out = {k: np.zeros((nx, ny, nz), order="F") for k in
           ("rhoinv", "c2rho0", "gradp0_z", "gradrho0_z", "az", "bzpml")}
    un = np.zeros(nz)
    st = np.zeros(nz)
    lib().orc_init_rhs(nx, ny, nz, int(is2d), npmlbot, npmltop, diml, dimc, _dp(z), _dp(c2),
                       _dp(rho0), _dp(p0), _dp(g), _dp(out["rhoinv"]), _dp(out["c2rho0"]),
                       _dp(out["gradp0_z"]), _dp(out["gradrho0_z"]), _dp(out["az"]),
                       _dp(out["bzpml"]), _dp(un), _dp(st))
    out["g"] = g
    out["unstretch"] = un
    out["stretch"] = st
    assert N == c2.size
    return out


def omp_threads() -> int:
    """OpenMP threads the oracle's loops use."""
    return int(lib().orc_threads())


class Oracle:
    """Stateful FP64 solver with the same set-up / step / get calls as the
    GPU engine's C ABI (include/sparc_gpu.h)."""

    def __init__(self, cfg):
        self.L = lib()
        self.cfg = cfg
        self.h = self.L.orc_create(C.byref(cfg))
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        self.shape = (cfg.nx, cfg.ny, cfg.nz)
        self.N = cfg.nx * cfg.ny * cfg.nz
        self.NP = cfg.nx * cfg.ny * max(cfg.npmlbot + cfg.npmltop, 1)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def beta(self, s):
        return self.L.orc_beta(self.h, s)

    def set_background(self, bid, arr):
        a = np.asfortranarray(arr, dtype=np.float64).ravel(order="F")
        assert self.L.orc_set_background(self.h, bid, _dp(a), a.size) == 0

    def set_profile(self, pid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert self.L.orc_set_profile(self.h, pid, _dp(a), a.size) == 0

    def set_damping(self, rates):
        a = np.asfortranarray(rates, dtype=np.float64).ravel(order="F")
        assert self.L.orc_set_damping(self.h, _dp(a), a.size) == 0

    def set_filter(self, dx, dy):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        assert self.L.orc_set_filter(self.h, _dp(dx), _dp(dy)) == 0

    def set_source(self, kind, mask, series, active_until):
        m = np.asfortranarray(mask, dtype=np.float64).ravel(order="F")
        s = np.ascontiguousarray(series if series is not None else np.zeros(1), dtype=np.float64)
        nt = s.size if kind == 1 else (m.size // (self.cfg.nx * self.cfg.ny))
        assert self.L.orc_set_source(self.h, kind, _dp(m), m.size, _dp(s), nt, active_until) == 0

    def set_source_planes(self, planes):
        if planes is None:
            assert self.L.orc_set_source_planes(self.h, None, 0) == 0
            return
        p = np.asfortranarray(planes, dtype=np.int32).ravel(order="F")
        assert self.L.orc_set_source_planes(self.h, p.ctypes.data_as(C.POINTER(C.c_int32)),
                                            p.size) == 0

    def step(self, t0, n):
        assert self.L.orc_step(self.h, t0, n) == 0

    def get(self, arr, field):
        n = self.N if arr in (0, 1) else self.NP
        out = np.zeros(n)
        assert self.L.orc_get_state(self.h, arr, field, _dp(out)) == 0
        if arr in (0, 1):
            return out.reshape(self.shape, order="F")
        return out

    def put(self, arr, field, data):
        a = np.asfortranarray(data, dtype=np.float64).ravel(order="F")
        assert self.L.orc_put_state(self.h, arr, field, _dp(a)) == 0

    def norm(self, field):
        return self.L.orc_norm(self.h, field)

    def damp_field(self, v, rhs):
        v = np.asfortranarray(v, dtype=np.float64)
        r = np.array(rhs, dtype=np.float64, order="F")
        self.L.orc_damp_field(self.h, _dp(v), _dp(r))
        return r

    def snapshots_alloc(self, n):
        assert self.L.orc_snapshots_alloc(self.h, n) == 0

    def snapshot_store(self, slot):
        assert self.L.orc_snapshot_store(self.h, slot) == 0

    def snapshot_get(self, slot, f):
        out = np.zeros(self.N)
        assert self.L.orc_snapshot_get(self.h, slot, f, _dp(out)) == 0
        return out.reshape(self.shape, order="F")

    def snapshot_put(self, slot, f, data):
        a = np.asfortranarray(data, dtype=np.float64).ravel(order="F")
        assert self.L.orc_snapshot_put(self.h, slot, f, _dp(a)) == 0

    def kernel_begin(self, grav):
        g = np.ascontiguousarray(grav, dtype=np.float64)
        assert self.L.orc_kernel_begin(self.h, _dp(g), g.size) == 0

    def kernel_accumulate(self, slot):
        assert self.L.orc_kernel_accumulate(self.h, slot) == 0

    def kernel_get(self, which):
        out = np.zeros(self.N)
        assert self.L.orc_kernel_get(self.h, which, _dp(out)) == 0
        return out.reshape(self.shape, order="F")

    def kernel_put(self, which, data):
        d = np.asfortranarray(data, dtype=np.float64)
        assert self.L.orc_kernel_put(self.h, which, _dp(d)) == 0
