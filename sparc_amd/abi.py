"""ctypes binding of the C ABI in include/sparc_gpu.h.

The engine library ``sparc_amd/libsparc_gpu.so`` is built in-tree by
``sparc_amd.build`` (hipcc, gfx950).  Loading fails loudly when it is absent:
there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparc_gpu.so")

ABI_VERSION = 1
OK, EINVAL, EHIP, ENOMEM, ESTATE = 0, -1, -2, -3, -4
FP64, FP32 = 0, 1
OPT_3D, OPT_2D = 4, 17
OPT_RHOVP = 2          # (rho, v, p) form, forward only
OPT_MHD = 10           # linear MHD, displacement form
OPT_MHD2D = 18         # linear MHD, displacement form, ny = 1
OPT_MHD8 = 8           # linear MHD, (rho, v, p, b) form, forward only
BX = 5                 # option 8 slot of b_x (b in 5..7)
RHO, P = 0, 4          # option 2 slots of rho and p (v in 1..3)
FORWARD, ADJOINT = 0, 1
XI_X, XI_Y, XI_Z, V_X, V_Y, V_Z = range(6)
ARR_A, ARR_SCR, ARR_PMLVARS, ARR_SCRPML, ARR_ALO = range(5)
(BG_RHO0, BG_RHOINV, BG_C2RHO0, BG_GRADP0_Z, BG_GRADRHO0_Z, BG_AZ, BG_BZPML,
 BG_C2, BG_BX, BG_BY, BG_BZ, BG_CURLBX, BG_CURLBY, BG_CURLBZ, BG_GRADP0_X, BG_GRADP0_Y,
 BG_GRADRHO0_X, BG_GRADRHO0_Y) = range(18)
PROF_STRETCH, PROF_UNSTRETCH, PROF_G, PROF_Z, PROF_SOURCE_DEP, PROF_SPONGE = range(6)
SRC_NONE, SRC_SERIES, SRC_LAGRANGE = range(3)
KERN_C2, KERN_RHO, KERN_HESSIAN, KERN_VX, KERN_VY, KERN_VZ, KERN_BX, KERN_BY, KERN_BZ, KERN_PSI = range(10)
(KC_XLINES, KC_YLINES, KC_ZSOLVE, KC_DAMP, KC_UPDATE, KC_OTHER, KC_ZSOLVE_P, KC_DAMP_Y,
 KC_DAMP_XI, KC_FILTER) = range(10)
KC_NAMES = ("xlines", "ylines", "zsolve", "damp", "update", "other", "zsolve_p", "damp_y",
            "damp_xi", "filter")
OP_DDX, OP_DDY, OP_DDZ, OP_ZFILTER, OP_DAMP, OP_LOWPASS_X, OP_LOWPASS_Y = range(7)

# every entry point declared in include/sparc_gpu.h
EXPORTS = (
    "sparc_create", "sparc_destroy", "sparc_last_error", "sparc_device_bytes",
    "sparc_footprint", "sparc_set_background", "sparc_set_profile", "sparc_set_damping",
    "sparc_set_filter", "sparc_set_source", "sparc_set_source_planes", "sparc_step", "sparc_sync", "sparc_reset",
    "sparc_get_state", "sparc_put_state", "sparc_get_plane", "sparc_norm", "sparc_sample",
    "sparc_sample_noalloc",
    "sparc_snapshots_alloc", "sparc_snapshot_store", "sparc_snapshots_copy",
    "sparc_kernel_begin", "sparc_kernel_accumulate", "sparc_kernel_get", "sparc_apply_op",
    "sparc_profile_enable", "sparc_profile_read", "sparc_kernel_accumulate_from",
    "sparc_checkpoints_alloc", "sparc_checkpoint_save", "sparc_checkpoint_load",
)


class SparcConfig(C.Structure):
    """Mirror of ``sparc_config`` (include/sparc_gpu.h)."""

    _fields_ = [
        ("abi_version", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("option", C.c_int32),
        ("precision", C.c_int32),
        ("npmlbot", C.c_int32),
        ("npmltop", C.c_int32),
        ("damp_waves", C.c_int32),
        ("freq_filtering", C.c_int32),
        ("e_rad", C.c_int32),
        ("o_rad", C.c_int32),
        ("mode", C.c_int32),
        ("cadforcing_step", C.c_int32),
        ("forcing_length", C.c_int32),
        ("num_steps", C.c_int32),
        ("compensated", C.c_int32),
        ("flow_kernels", C.c_int32),
        ("use_graphs", C.c_int32),
        ("plane_damping", C.c_int32),
        ("reserved_i", C.c_int32 * 4),
        ("xlength", C.c_double), ("ylength", C.c_double),
        ("timestep", C.c_double),
        ("deltat", C.c_double),
        ("stretchx", C.c_double), ("stretchy", C.c_double),
        ("delta_width", C.c_double),
        ("cadforcing", C.c_double),
        ("field_scale", C.c_double),
        ("reserved_d", C.c_double * 7),
    ]

    def copy(self) -> "SparcConfig":
        c = SparcConfig()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(SparcConfig))
        return c


DP = C.POINTER(C.c_double)
_LIB = None


def _dp(a: np.ndarray):
    return a.ctypes.data_as(DP)


def lib() -> C.CDLL:
    """Loads libsparc_gpu.so; raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP engine not built: {LIB_PATH} missing (run python -c "
            "'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sparc_create.argtypes = [C.POINTER(SparcConfig), C.POINTER(vp)]
    L.sparc_destroy.argtypes = [vp]
    L.sparc_destroy.restype = None
    L.sparc_last_error.argtypes = [vp]
    L.sparc_last_error.restype = C.c_char_p
    L.sparc_device_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.sparc_footprint.argtypes = [C.POINTER(SparcConfig), C.POINTER(C.c_size_t)]
    L.sparc_set_background.argtypes = [vp, C.c_int, DP, C.c_size_t]
    L.sparc_set_profile.argtypes = [vp, C.c_int, DP, C.c_int]
    L.sparc_set_damping.argtypes = [vp, DP, C.c_size_t]
    L.sparc_set_filter.argtypes = [vp, DP, DP]
    L.sparc_set_source.argtypes = [vp, C.c_int, DP, C.c_size_t, DP, C.c_int, C.c_int]
    L.sparc_set_source_planes.argtypes = [vp, C.POINTER(C.c_int32), C.c_size_t]
    L.sparc_step.argtypes = [vp, C.c_int, C.c_int]
    L.sparc_sync.argtypes = [vp]
    L.sparc_reset.argtypes = [vp]
    L.sparc_get_state.argtypes = [vp, C.c_int, C.c_int, DP]
    L.sparc_put_state.argtypes = [vp, C.c_int, C.c_int, DP]
    L.sparc_get_plane.argtypes = [vp, C.c_int, C.c_int, DP]
    L.sparc_norm.argtypes = [vp, C.c_int, DP]
    L.sparc_sample.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int32), DP]
    L.sparc_sample_noalloc.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int32), DP]
    L.sparc_snapshots_alloc.argtypes = [vp, C.c_int]
    L.sparc_snapshot_store.argtypes = [vp, C.c_int]
    L.sparc_snapshots_copy.argtypes = [vp, vp]
    L.sparc_kernel_begin.argtypes = [vp, DP, C.c_int]
    L.sparc_kernel_accumulate.argtypes = [vp, C.c_int]
    L.sparc_kernel_get.argtypes = [vp, C.c_int, DP]
    L.sparc_apply_op.argtypes = [vp, C.c_int, C.c_int, DP, DP]
    L.sparc_kernel_accumulate_from.argtypes = [vp, vp, C.c_int]
    L.sparc_checkpoints_alloc.argtypes = [vp, C.c_int]
    L.sparc_checkpoint_save.argtypes = [vp, C.c_int]
    L.sparc_checkpoint_load.argtypes = [vp, C.c_int]
    L.sparc_profile_enable.argtypes = [vp, C.c_int]
    L.sparc_profile_read.argtypes = [vp, C.c_int, DP, DP, C.POINTER(C.c_int64)]
    _LIB = L
    return L


class SparcError(RuntimeError):
    pass


def footprint(cfg: SparcConfig) -> int:
    n = C.c_size_t(0)
    rc = lib().sparc_footprint(C.byref(cfg), C.byref(n))
    if rc != OK:
        raise SparcError(f"sparc_footprint rejected the configuration ({rc})")
    return int(n.value)


class Engine:
    """One device context: the state and work arrays of one solve."""

    def __init__(self, cfg: SparcConfig):
        self.L = lib()
        self.cfg = cfg.copy()
        self.h = C.c_void_p()
        rc = self.L.sparc_create(C.byref(self.cfg), C.byref(self.h))
        if rc != OK:
            msg = self.L.sparc_last_error(self.h).decode() if self.h else ""
            if self.h:
                self.L.sparc_destroy(self.h)
                self.h = C.c_void_p()
            raise SparcError(f"sparc_create failed ({rc}): {msg}")
        self.shape = (cfg.nx, cfg.ny, cfg.nz)
        self.N = cfg.nx * cfg.ny * cfg.nz
        self.NP = cfg.nx * cfg.ny * max(cfg.npmlbot + cfg.npmltop, 1)

    def _chk(self, rc, what):
        if rc != OK:
            raise SparcError(f"{what} failed ({rc}): {self.L.sparc_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.sparc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        n = C.c_size_t(0)
        self._chk(self.L.sparc_device_bytes(self.h, C.byref(n)), "sparc_device_bytes")
        return int(n.value)

    def set_background(self, bid: int, arr):
        a = np.asfortranarray(arr, dtype=np.float64).ravel(order="F")
        self._chk(self.L.sparc_set_background(self.h, bid, _dp(a), a.size), "set_background")

    def set_profile(self, pid: int, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.sparc_set_profile(self.h, pid, _dp(a), a.size), "set_profile")

    def set_damping(self, rates):
        a = np.asfortranarray(rates, dtype=np.float64).ravel(order="F")
        self._chk(self.L.sparc_set_damping(self.h, _dp(a), a.size), "set_damping")

    def set_filter(self, dx, dy):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dy = np.ascontiguousarray(dy, dtype=np.float64)
        self._chk(self.L.sparc_set_filter(self.h, _dp(dx), _dp(dy)), "set_filter")

    def set_source(self, kind: int, mask, series, active_until: int):
        if kind == SRC_NONE:
            self._chk(self.L.sparc_set_source(self.h, SRC_NONE, None, 0, None, 0, 0), "set_source")
            return
        m = np.asfortranarray(mask, dtype=np.float64).ravel(order="F")
        if kind == SRC_SERIES:
            s = np.ascontiguousarray(series, dtype=np.float64)
            nt = s.size
        else:
            s = np.zeros(1)
            nt = m.size // (self.cfg.nx * self.cfg.ny)
        self._chk(self.L.sparc_set_source(self.h, kind, _dp(m), m.size, _dp(s), nt, active_until),
                  "set_source")

    def set_source_planes(self, planes):
        """1-based per-pixel source planes (nx, ny), or None for the uniform plane."""
        if planes is None:
            self._chk(self.L.sparc_set_source_planes(self.h, None, 0), "set_source_planes")
            return
        p = np.asfortranarray(planes, dtype=np.int32).ravel(order="F")
        self._chk(self.L.sparc_set_source_planes(
            self.h, p.ctypes.data_as(C.POINTER(C.c_int32)), p.size), "set_source_planes")

    def step(self, t0: int, n: int):
        self._chk(self.L.sparc_step(self.h, t0, n), "sparc_step")

    def sync(self):
        self._chk(self.L.sparc_sync(self.h), "sparc_sync")

    def reset(self):
        self._chk(self.L.sparc_reset(self.h), "sparc_reset")

    def get(self, arr: int, field: int) -> np.ndarray:
        full = arr in (ARR_A, ARR_SCR, ARR_ALO)
        out = np.empty(self.N if full else self.NP)
        self._chk(self.L.sparc_get_state(self.h, arr, field, _dp(out)), "get_state")
        return out.reshape(self.shape, order="F") if full else out

    def put(self, arr: int, field: int, data):
        a = np.asfortranarray(data, dtype=np.float64).ravel(order="F")
        self._chk(self.L.sparc_put_state(self.h, arr, field, _dp(a)), "put_state")

    def plane(self, field: int, k: int) -> np.ndarray:
        out = np.empty(self.cfg.nx * self.cfg.ny)
        self._chk(self.L.sparc_get_plane(self.h, field, k, _dp(out)), "get_plane")
        return out.reshape((self.cfg.nx, self.cfg.ny), order="F")

    def norm(self, field: int) -> float:
        v = C.c_double(0)
        self._chk(self.L.sparc_norm(self.h, field, C.byref(v)), "sparc_norm")
        return v.value

    def sample(self, field: int, ijk: np.ndarray, noalloc: bool = False) -> np.ndarray:
        """a(i,j,k,field) at the points; noalloc stages them in the library's
        fixed device array (sparc_sample_noalloc) instead of allocating."""
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        out = np.empty(ijk.shape[0])
        fn = self.L.sparc_sample_noalloc if noalloc else self.L.sparc_sample
        self._chk(fn(self.h, field, ijk.shape[0], ijk.ctypes.data_as(C.POINTER(C.c_int32)),
                     _dp(out)), "sparc_sample_noalloc" if noalloc else "sparc_sample")
        return out

    def snapshots_alloc(self, n: int):
        self._chk(self.L.sparc_snapshots_alloc(self.h, n), "snapshots_alloc")

    def snapshot_store(self, slot: int):
        self._chk(self.L.sparc_snapshot_store(self.h, slot), "snapshot_store")

    def snapshots_copy_from(self, other: "Engine"):
        self._chk(self.L.sparc_snapshots_copy(self.h, other.h), "snapshots_copy")

    def kernel_begin(self, grav):
        g = np.ascontiguousarray(grav, dtype=np.float64)
        self._chk(self.L.sparc_kernel_begin(self.h, _dp(g), g.size), "kernel_begin")

    def kernel_accumulate(self, slot: int):
        self._chk(self.L.sparc_kernel_accumulate(self.h, slot), "kernel_accumulate")

    def apply_op(self, op: int, f, ibc: int = 1, out=None) -> np.ndarray:
        """One operator through the stepper's kernels; ``out`` carries the
        caller boundary planes of dbyd2 for ibc 2/3."""
        a = np.asfortranarray(f, dtype=np.float64).ravel(order="F")
        o = (np.zeros(self.N) if out is None
             else np.array(out, dtype=np.float64, order="F").ravel(order="F"))
        self._chk(self.L.sparc_apply_op(self.h, op, ibc, _dp(a), _dp(o)), "apply_op")
        return o.reshape(self.shape, order="F")

    def kernel_accumulate_from(self, src: "Engine", slot: int):
        self._chk(self.L.sparc_kernel_accumulate_from(self.h, src.h, slot), "kernel_accumulate_from")

    def checkpoints_alloc(self, n: int):
        self._chk(self.L.sparc_checkpoints_alloc(self.h, n), "checkpoints_alloc")

    def checkpoint_save(self, slot: int):
        self._chk(self.L.sparc_checkpoint_save(self.h, slot), "checkpoint_save")

    def checkpoint_load(self, slot: int):
        self._chk(self.L.sparc_checkpoint_load(self.h, slot), "checkpoint_load")

    def profile(self, on: bool):
        self._chk(self.L.sparc_profile_enable(self.h, int(on)), "profile_enable")

    def profile_read(self) -> dict:
        """kernel class -> (ms, algorithmic bytes, launches)."""
        out = {}
        for k, name in enumerate(KC_NAMES):
            ms, by, n = C.c_double(0), C.c_double(0), C.c_int64(0)
            self._chk(self.L.sparc_profile_read(self.h, k, C.byref(ms), C.byref(by), C.byref(n)),
                      "profile_read")
            out[name] = (ms.value, by.value, int(n.value))
        return out

    def kernel_get(self, which: int) -> np.ndarray:
        out = np.empty(self.N)
        self._chk(self.L.sparc_kernel_get(self.h, which, _dp(out)), "kernel_get")
        return out.reshape(self.shape, order="F")
