"""ctypes binding of libclue_hip.so (the C ABI in include/clue_hip.h).

Device memory is managed through the library's own helpers, so a Python caller needs nothing but
numpy. Loading fails loudly when the library is not built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libclue_hip.so"

CLUE_OK = 0
ERRORS = {-1: "CLUE_ERR_ARG", -2: "CLUE_ERR_HIP", -3: "CLUE_ERR_BUDGET", -4: "CLUE_ERR_LAYER", -5: "CLUE_ERR_BIN",
          -6: "CLUE_ERR_TIMEOUT", -7: "CLUE_ERR_CAPACITY"}
PATH_AUTO, PATH_FUSED, PATH_GLOBAL = 0, 1, 2
PLAN_GRAPH, PLAN_AUTOTUNE, PLAN_INDEX_PASS = 1, 2, 4
FUSED_CAPACITY = 10112
NOALLOC_MAX_POINTS = 1048576

# Every symbol include/clue_hip.h declares.
EXPORTS = [
    "clue3d_default_params", "clue_create", "clue_destroy", "clue_last_error", "clue_device_count",
    "clue_bytes_in_use", "clue_stream_create", "clue_stream_destroy", "clue_stream_sync", "clue_stream_callback",
    "clue_malloc_device", "clue_cache_alloc", "clue_cache_free", "clue_cache_release", "clue_cache_stats",
    "clue_free_device", "clue_malloc_pinned", "clue_free_pinned", "clue_memcpy_htod", "clue_memcpy_dtoh",
    "clue_memcpy_dtod",
    "clue_check", "clue_mark", "clue_elapsed_ms", "clue2d_run_device", "clue2d_run_host",
    "clue2d_run_batch_device", "clue2d_plan_create", "clue3d_run_device", "clue3d_run_host",
    "clue3d_run_batch_device", "clue3d_plan_create", "clue_plan_launch", "clue_plan_destroy",
    "clue_plan_workspace_bytes", "clue_selftest_arith", "clue_server_create", "clue_server_destroy",
    "clue_server_submit2d", "clue_server_submit3d", "clue_server_wait", "clue_server_stats",
    "clue_server_last_error", "clue_set_path", "clue_plan_path", "clue_plan_tuning", "clue2d_run_device_noalloc",
    "clue_plan_kernel_ms", "clue_plan_index_pass",
]


class Clue2DParams(C.Structure):
    _fields_ = [("dc", C.c_float), ("rhoc", C.c_float), ("outlierDeltaFactor", C.c_float)]


class Clue3DParams(C.Structure):
    _fields_ = [
        ("densitySiblingLayers", C.c_int32), ("densityXYDistanceSqr", C.c_float),
        ("kernelDensityFactor", C.c_float), ("densityOnSameLayer", C.c_int32),
        ("nearestHigherOnSameLayer", C.c_int32), ("criticalXYDistance", C.c_float),
        ("criticalZDistanceLyr", C.c_float), ("criticalDensity", C.c_float),
        ("criticalSelfDensity", C.c_float), ("outlierMultiplier", C.c_float),
    ]


VP = C.c_void_p
I32 = C.c_int32

FIELDS_2D_IN = ("x", "y", "layer", "weight")
FIELDS_2D_OUT = ("rho", "delta", "nearestHigher", "isSeed", "clusterIndex")
FIELDS_3D_IN = ("z", "eta", "phi", "r_over_absz", "radius", "energy", "layer", "isSilicon")
FIELDS_3D_OUT = ("rho", "delta_f", "delta_l", "nh_layer", "nh_idx", "isSeed", "tracksterIndex")
_INT_FIELDS = {"layer", "isSilicon", "nearestHigher", "isSeed", "clusterIndex", "delta_l", "nh_layer", "nh_idx",
               "tracksterIndex"}


def _dt(name):
    return np.int32 if name in _INT_FIELDS else np.float32


def load_library(path: Path = LIB_PATH) -> C.CDLL:
    if not Path(path).exists():
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (make -C {HERE})")
    lib = C.CDLL(str(path))
    sig = {
        "clue3d_default_params": (Clue3DParams, []),
        "clue_create": (VP, [I32, C.c_uint64]),
        "clue_destroy": (None, [VP]),
        "clue_last_error": (C.c_char_p, [VP]),
        "clue_device_count": (I32, []),
        "clue_bytes_in_use": (C.c_uint64, [VP]),
        "clue_stream_create": (VP, [VP]),
        "clue_stream_destroy": (I32, [VP, VP]),
        "clue_stream_sync": (I32, [VP, VP]),
        "clue_stream_callback": (I32, [VP, VP, VP, VP]),
        "clue_cache_alloc": (VP, [VP, C.c_uint64, VP]),
        "clue_cache_free": (I32, [VP, VP]),
        "clue_cache_release": (I32, [VP]),
        "clue_cache_stats": (I32, [VP, VP, VP, VP, VP]),
        "clue_malloc_device": (VP, [VP, C.c_uint64]),
        "clue_free_device": (I32, [VP, VP]),
        "clue_malloc_pinned": (VP, [VP, C.c_uint64]),
        "clue_free_pinned": (I32, [VP, VP]),
        "clue_memcpy_htod": (I32, [VP, VP, VP, C.c_uint64, VP]),
        "clue_memcpy_dtoh": (I32, [VP, VP, VP, C.c_uint64, VP]),
        "clue_memcpy_dtod": (I32, [VP, VP, VP, C.c_uint64, VP]),
        "clue_check": (I32, [VP, VP]),
        "clue_mark": (I32, [VP, I32, VP]),
        "clue_elapsed_ms": (C.c_float, [VP, I32, I32]),
        "clue2d_run_device": (I32, [VP, C.POINTER(Clue2DParams), I32] + [VP] * 9 + [VP]),
        "clue2d_run_device_noalloc": (I32, [VP, C.POINTER(Clue2DParams), I32] + [VP] * 9 + [VP]),
        "clue2d_run_host": (I32, [VP, C.POINTER(Clue2DParams), I32] + [VP] * 9),
        "clue2d_run_batch_device": (I32, [VP, C.POINTER(Clue2DParams), I32, VP] + [VP] * 9 + [VP]),
        "clue2d_plan_create": (VP, [VP, C.POINTER(Clue2DParams), I32, VP] + [VP] * 9 + [I32]),
        "clue3d_run_device": (I32, [VP, C.POINTER(Clue3DParams), I32] + [VP] * 15 + [VP]),
        "clue3d_run_host": (I32, [VP, C.POINTER(Clue3DParams), I32] + [VP] * 15),
        "clue3d_run_batch_device": (I32, [VP, C.POINTER(Clue3DParams), I32, VP] + [VP] * 15 + [VP]),
        "clue3d_plan_create": (VP, [VP, C.POINTER(Clue3DParams), I32, VP] + [VP] * 15 + [I32]),
        "clue_plan_launch": (I32, [VP, VP]),
        "clue_plan_destroy": (None, [VP]),
        "clue_plan_workspace_bytes": (C.c_uint64, [VP]),
        "clue_selftest_arith": (I32, [VP, I32, VP, VP] + [VP] * 7),
        "clue_server_create": (VP, [VP, C.POINTER(Clue2DParams), C.POINTER(Clue3DParams), C.c_int64, C.c_int64,
                                    I32, I32]),
        "clue_server_destroy": (None, [VP]),
        "clue_server_submit2d": (C.c_int64, [VP, I32] + [VP] * 9),
        "clue_server_submit3d": (C.c_int64, [VP, I32] + [VP] * 15),
        "clue_server_wait": (I32, [VP, C.c_int64, I32, C.POINTER(C.c_double)]),
        "clue_server_stats": (I32, [VP] + [C.POINTER(C.c_int64)] * 3),
        "clue_server_last_error": (C.c_char_p, [VP]),
        "clue_set_path": (I32, [VP, I32, I32]),
        "clue_plan_path": (I32, [VP]),
        "clue_plan_index_pass": (I32, [VP]),
        "clue_plan_tuning": (I32, [VP, VP]),
        "clue_plan_kernel_ms": (C.c_float, [VP, I32, I32, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def params2d(dc=3.0, rhoc=10.0, odf=2.0) -> Clue2DParams:
    return Clue2DParams(dc, rhoc, odf)


def params3d(S=3, **kw) -> Clue3DParams:
    d = dict(densitySiblingLayers=S, densityXYDistanceSqr=3.24, kernelDensityFactor=0.2, densityOnSameLayer=0,
             nearestHigherOnSameLayer=0, criticalXYDistance=1.8, criticalZDistanceLyr=5.0, criticalDensity=0.6,
             criticalSelfDensity=0.15, outlierMultiplier=2.0)
    d.update(kw)
    return Clue3DParams(**d)


class ClueError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class DeviceArray:
    def __init__(self, dev: "Device", n: int, dtype):
        self.dev, self.n, self.dtype = dev, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = dev.lib.clue_malloc_device(dev.ctx, max(self.nbytes, 4))
        if not self.ptr:
            raise ClueError(-3, dev.error())

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        self.dev._chk(self.dev.lib.clue_memcpy_htod(self.dev.ctx, self.ptr, a.ctypes.data, a.nbytes, stream))
        if stream is None:
            self.dev.sync()
        return self

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        self.dev._chk(self.dev.lib.clue_memcpy_dtoh(self.dev.ctx, out.ctypes.data, self.ptr, self.nbytes, stream))
        self.dev.sync(stream)
        return out

    def free(self):
        if self.ptr:
            self.dev.lib.clue_free_device(self.dev.ctx, self.ptr)
            self.ptr = None


class Device:
    """One clue_ctx on one GPU."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, lib: C.CDLL | None = None):
        self.lib = lib or load_library()
        if self.lib.clue_device_count() <= device:
            raise RuntimeError("no HIP device available for libclue_hip")
        self.ctx = self.lib.clue_create(device, budget_bytes)
        if not self.ctx:
            raise RuntimeError("clue_create failed")

    def close(self):
        if self.ctx:
            for ptr in getattr(self, "_pinned", []):
                self.lib.clue_free_pinned(self.ctx, ptr)
            self._pinned = []
            self.lib.clue_destroy(self.ctx)
            self.ctx = None

    def pinned_copy(self, a: np.ndarray) -> np.ndarray:
        """A numpy array in pinned host memory (freed with the device) holding a copy of `a`."""
        a = np.ascontiguousarray(a)
        ptr = self.lib.clue_malloc_pinned(self.ctx, max(a.nbytes, 4))
        if not ptr:
            raise ClueError(-2, self.error())
        if not hasattr(self, "_pinned"):
            self._pinned = []
        self._pinned.append(ptr)
        buf = (C.c_char * max(a.nbytes, 4)).from_address(ptr)
        out = np.frombuffer(buf, dtype=a.dtype, count=a.size)
        out[...] = a
        return out

    def error(self) -> str:
        return self.lib.clue_last_error(self.ctx).decode()

    def _chk(self, rc):
        if rc != CLUE_OK:
            raise ClueError(rc, self.error())

    def sync(self, stream=None):
        self._chk(self.lib.clue_stream_sync(self.ctx, stream))

    def check(self, stream=None):
        self._chk(self.lib.clue_check(self.ctx, stream))

    def array(self, n, dtype) -> DeviceArray:
        return DeviceArray(self, n, dtype)

    def upload(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray(self, a.size, a.dtype).upload(a)

    def upload_soa(self, data: dict, fields) -> dict:
        return {k: self.upload(np.ascontiguousarray(data[k], dtype=_dt(k))) for k in fields}

    def alloc_soa(self, n: int, fields) -> dict:
        return {k: self.array(n, _dt(k)) for k in fields}

    # ------------------------------------------------------------ CLUE2D
    def set_path(self, path: int, dim: int = 2):
        """CLUE2D execution path for later calls and plans: PATH_AUTO, PATH_FUSED or PATH_GLOBAL."""
        self._chk(self.lib.clue_set_path(self.ctx, dim, path))

    def clue2d_host(self, ev: dict, p: Clue2DParams | None = None) -> dict:
        p = p or params2d()
        n = ev["x"].size
        ins = [np.ascontiguousarray(ev[k], dtype=_dt(k)) for k in FIELDS_2D_IN]
        outs = {k: np.empty(n, _dt(k)) for k in FIELDS_2D_OUT}
        self._chk(self.lib.clue2d_run_host(self.ctx, C.byref(p), n, *[a.ctypes.data for a in ins],
                                           *[outs[k].ctypes.data for k in FIELDS_2D_OUT]))
        return outs

    def clue2d_device(self, din: dict, dout: dict, n: int, p: Clue2DParams | None = None, stream=None):
        p = p or params2d()
        self._chk(self.lib.clue2d_run_device(self.ctx, C.byref(p), n, *[din[k].ptr for k in FIELDS_2D_IN],
                                             *[dout[k].ptr for k in FIELDS_2D_OUT], stream))

    def clue2d_device_noalloc(self, din: dict, dout: dict, n: int, p: Clue2DParams | None = None, stream=None):
        """clue2d_run_device on the library's fixed scratch (no allocation; n <= NOALLOC_MAX_POINTS)."""
        p = p or params2d()
        self._chk(self.lib.clue2d_run_device_noalloc(self.ctx, C.byref(p), n, *[din[k].ptr for k in FIELDS_2D_IN],
                                                     *[dout[k].ptr for k in FIELDS_2D_OUT], stream))

    def clue2d_batch(self, din: dict, dout: dict, offs: np.ndarray, p: Clue2DParams | None = None, stream=None):
        p = p or params2d()
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        self._chk(self.lib.clue2d_run_batch_device(self.ctx, C.byref(p), offs.size - 1, offs.ctypes.data,
                                                   *[din[k].ptr for k in FIELDS_2D_IN],
                                                   *[dout[k].ptr for k in FIELDS_2D_OUT], stream))

    def clue2d_plan(self, din: dict, dout: dict, offs: np.ndarray, p: Clue2DParams | None = None,
                    use_graph: bool = True, autotune: bool = False, index_pass: bool = False) -> "Plan":
        p = p or params2d()
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        flags = ((PLAN_GRAPH if use_graph else 0) | (PLAN_AUTOTUNE if autotune else 0)
                 | (PLAN_INDEX_PASS if index_pass else 0))
        h = self.lib.clue2d_plan_create(self.ctx, C.byref(p), offs.size - 1, offs.ctypes.data,
                                        *[din[k].ptr for k in FIELDS_2D_IN], *[dout[k].ptr for k in FIELDS_2D_OUT],
                                        flags)
        if not h:
            raise ClueError(-1, self.error())
        return Plan(self, h)

    # ------------------------------------------------------------ CLUE3D
    def clue3d_host(self, ev: dict, p: Clue3DParams | None = None) -> dict:
        p = p or params3d()
        n = ev["layer"].size
        ins = [np.ascontiguousarray(ev[k], dtype=_dt(k)) for k in FIELDS_3D_IN]
        outs = {k: np.empty(n, _dt(k)) for k in FIELDS_3D_OUT}
        self._chk(self.lib.clue3d_run_host(self.ctx, C.byref(p), n, *[a.ctypes.data for a in ins],
                                           *[outs[k].ctypes.data for k in FIELDS_3D_OUT]))
        return outs

    def clue3d_device(self, din: dict, dout: dict, n: int, p: Clue3DParams | None = None, stream=None):
        p = p or params3d()
        self._chk(self.lib.clue3d_run_device(self.ctx, C.byref(p), n, *[din[k].ptr for k in FIELDS_3D_IN],
                                             *[dout[k].ptr for k in FIELDS_3D_OUT], stream))

    def clue3d_batch(self, din: dict, dout: dict, offs: np.ndarray, p: Clue3DParams | None = None, stream=None):
        p = p or params3d()
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        self._chk(self.lib.clue3d_run_batch_device(self.ctx, C.byref(p), offs.size - 1, offs.ctypes.data,
                                                   *[din[k].ptr for k in FIELDS_3D_IN],
                                                   *[dout[k].ptr for k in FIELDS_3D_OUT], stream))

    def clue3d_plan(self, din: dict, dout: dict, offs: np.ndarray, p: Clue3DParams | None = None,
                    use_graph: bool = True, autotune: bool = False) -> "Plan":
        p = p or params3d()
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        flags = (PLAN_GRAPH if use_graph else 0) | (PLAN_AUTOTUNE if autotune else 0)
        h = self.lib.clue3d_plan_create(self.ctx, C.byref(p), offs.size - 1, offs.ctypes.data,
                                        *[din[k].ptr for k in FIELDS_3D_IN], *[dout[k].ptr for k in FIELDS_3D_OUT],
                                        flags)
        if not h:
            raise ClueError(-1, self.error())
        return Plan(self, h)

    # ------------------------------------------------------------ helpers
    @staticmethod
    def download_soa(d: dict) -> dict:
        return {k: v.download() for k, v in d.items()}

    def selftest_arith(self, v: np.ndarray, w: np.ndarray) -> dict:
        v = np.ascontiguousarray(v, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        n = v.size
        o = {k: np.empty(n, np.int32) for k in ("xbin", "ybin", "eta_bin", "phi_bin")}
        o.update({k: np.empty(n, np.float32) for k in ("reduce", "dist", "quot")})
        self._chk(self.lib.clue_selftest_arith(self.ctx, n, v.ctypes.data, w.ctypes.data,
                                               *[o[k].ctypes.data for k in ("xbin", "ybin", "eta_bin", "phi_bin",
                                                                            "reduce", "dist", "quot")]))
        return o


class Server:
    """Continuous-batching request loop (clue_server_*). Submitted numpy arrays are kept referenced
    until the request is waited for."""

    def __init__(self, dev: "Device", p2: Clue2DParams | None = None, p3: Clue3DParams | None = None,
                 max_points_2d: int = 4_000_000, max_points_3d: int = 1_000_000, max_events: int = 64,
                 slots: int = 3):
        self.dev = dev
        self._p2, self._p3 = p2 or params2d(), p3 or params3d()
        self.h = dev.lib.clue_server_create(dev.ctx, C.byref(self._p2), C.byref(self._p3), max_points_2d,
                                            max_points_3d, max_events, slots)
        if not self.h:
            raise ClueError(-2, dev.error())
        self._live = {}

    def _err(self):
        return self.dev.lib.clue_server_last_error(self.h).decode()

    def submit2d(self, ev: dict, outs: dict | None = None) -> int:
        """Queue one CLUE2D event; `outs` may give the output arrays (e.g. pinned ones)."""
        ins = [np.ascontiguousarray(ev[k], dtype=_dt(k)) for k in FIELDS_2D_IN]
        n = ins[0].size
        outs = outs if outs is not None else {k: np.empty(n, _dt(k)) for k in FIELDS_2D_OUT}
        assert all(outs[k].size == n and outs[k].dtype == _dt(k) and outs[k].flags.c_contiguous for k in FIELDS_2D_OUT)
        rid = self.dev.lib.clue_server_submit2d(self.h, n, *[a.ctypes.data for a in ins],
                                                *[outs[k].ctypes.data for k in FIELDS_2D_OUT])
        if rid < 0:
            raise ClueError(int(rid), self._err())
        self._live[rid] = (ins, outs)
        return int(rid)

    def submit3d(self, ev: dict, outs: dict | None = None) -> int:
        """Queue one CLUE3D event; `outs` may give the output arrays (e.g. pinned ones)."""
        ins = [np.ascontiguousarray(ev[k], dtype=_dt(k)) for k in FIELDS_3D_IN]
        n = ins[0].size
        outs = outs if outs is not None else {k: np.empty(n, _dt(k)) for k in FIELDS_3D_OUT}
        assert all(outs[k].size == n and outs[k].dtype == _dt(k) and outs[k].flags.c_contiguous for k in FIELDS_3D_OUT)
        rid = self.dev.lib.clue_server_submit3d(self.h, n, *[a.ctypes.data for a in ins],
                                                *[outs[k].ctypes.data for k in FIELDS_3D_OUT])
        if rid < 0:
            raise ClueError(int(rid), self._err())
        self._live[rid] = (ins, outs)
        return int(rid)

    def wait(self, rid: int, timeout_ms: int = 60_000):
        lat = C.c_double(0.0)
        rc = self.dev.lib.clue_server_wait(self.h, rid, timeout_ms, C.byref(lat))
        if rc != CLUE_OK:
            raise ClueError(rc, self._err())
        _, outs = self._live.pop(rid)
        return outs, lat.value

    def outputs(self, rid: int) -> dict:
        """The output arrays of a request not yet waited for (kept referenced by the server object)."""
        return self._live[rid][1]

    def stats(self):
        b, e, p = C.c_int64(), C.c_int64(), C.c_int64()
        self.dev.lib.clue_server_stats(self.h, C.byref(b), C.byref(e), C.byref(p))
        return {"batches": b.value, "events": e.value, "points": p.value}

    def close(self):
        if self.h:
            self.dev.lib.clue_server_destroy(self.h)
            self.h = None


class Plan:
    def __init__(self, dev: Device, handle):
        self.dev, self.h = dev, handle

    def launch(self, stream=None):
        self.dev._chk(self.dev.lib.clue_plan_launch(self.h, stream))

    def kernel_ms(self, which: int, reps: int = 10, stream=None) -> float:
        """Mean ms of the plan's density (0) or nearest-higher (1) kernel, timed with HIP events."""
        ms = float(self.dev.lib.clue_plan_kernel_ms(self.h, which, reps, stream))
        if ms < 0:
            raise ClueError(-1, "clue_plan_kernel_ms failed")
        return ms

    @property
    def workspace_bytes(self) -> int:
        return int(self.dev.lib.clue_plan_workspace_bytes(self.h))

    @property
    def tuning(self) -> dict:
        """Points per workgroup and staged points of the density and nearest-higher kernels."""
        out = np.zeros(4, np.int32)
        self.dev._chk(self.dev.lib.clue_plan_tuning(self.h, out.ctypes.data))
        return {"density_points": int(out[0]), "density_stage": int(out[1]), "nh_points": int(out[2]),
                "nh_stage": int(out[3]), "nh_index_pass": self.index_pass}

    @property
    def index_pass(self) -> bool:
        """CLUE2D: whether the nearest-higher step uses the index-pass output form."""
        return bool(self.dev.lib.clue_plan_index_pass(self.h))

    @property
    def path(self) -> int:
        """PATH_FUSED or PATH_GLOBAL (CLUE2D: -1 until the first launch has chosen)."""
        return int(self.dev.lib.clue_plan_path(self.h))

    def destroy(self):
        if self.h:
            self.dev.lib.clue_plan_destroy(self.h)
            self.h = None
