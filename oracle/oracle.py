"""ctypes bindings of the CPU checkers. TEST INFRASTRUCTURE ONLY.

`Oracle` wraps liboracle.so (the C restatement); `Reference` wraps _ref/libclue_ref.so (the
reference's own serial kernels compiled from its sources). Only tests, the smoke check and the
CPU-baseline leg of the benchmark use this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libclue_ref.so"

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class Clue3DParams(C.Structure):
    _fields_ = [
        ("densitySiblingLayers", C.c_int32),
        ("densityXYDistanceSqr", C.c_float),
        ("kernelDensityFactor", C.c_float),
        ("densityOnSameLayer", C.c_int32),
        ("nearestHigherOnSameLayer", C.c_int32),
        ("criticalXYDistance", C.c_float),
        ("criticalZDistanceLyr", C.c_float),
        ("criticalDensity", C.c_float),
        ("criticalSelfDensity", C.c_float),
        ("outlierMultiplier", C.c_float),
    ]


def clue3d_params(S=3, **kw) -> Clue3DParams:
    """Reference defaults (CLUE3DAlgoKernels.h:26-32,286-290,509-515); S overrides siblings."""
    d = dict(densitySiblingLayers=S, densityXYDistanceSqr=3.24, kernelDensityFactor=0.2,
             densityOnSameLayer=0, nearestHigherOnSameLayer=0, criticalXYDistance=1.8,
             criticalZDistanceLyr=5.0, criticalDensity=0.6, criticalSelfDensity=0.15,
             outlierMultiplier=2.0)
    d.update(kw)
    return Clue3DParams(**d)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _out2d(n):
    return dict(rho=np.zeros(n, np.float32), delta=np.zeros(n, np.float32),
                nearestHigher=np.zeros(n, np.int32), isSeed=np.zeros(n, np.int32),
                clusterIndex=np.zeros(n, np.int32))


def _out3d(n):
    return dict(rho=np.zeros(n, np.float32), delta_f=np.zeros(n, np.float32), delta_l=np.zeros(n, np.int32),
                nh_layer=np.zeros(n, np.int32), nh_idx=np.zeros(n, np.int32), isSeed=np.zeros(n, np.int32),
                tracksterIndex=np.zeros(n, np.int32))


_ARGS2D = [C.c_int32, F32P, F32P, I32P, F32P, C.c_float, C.c_float, C.c_float, F32P, F32P, I32P, I32P, I32P]
_ARGS3D = [C.c_int32, F32P, F32P, F32P, F32P, F32P, F32P, I32P, I32P, C.POINTER(Clue3DParams),
           F32P, F32P, I32P, I32P, I32P, I32P, I32P]


class _Base:
    prefix = ""

    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run `make -C {HERE}`)")
        self.lib = C.CDLL(str(path))
        f = getattr(self.lib, self.prefix + "clue2d")
        f.argtypes, f.restype = _ARGS2D, C.c_int
        f = getattr(self.lib, self.prefix + "clue3d")
        f.argtypes, f.restype = _ARGS3D, C.c_int

    def clue2d(self, x, y, layer, weight, dc=3.0, rhoc=10.0, odf=2.0) -> dict:
        x, y, w = _c(x, np.float32), _c(y, np.float32), _c(weight, np.float32)
        layer = _c(layer, np.int32)
        n = x.size
        o = _out2d(n)
        rc = getattr(self.lib, self.prefix + "clue2d")(n, x, y, layer, w, dc, rhoc, odf, o["rho"], o["delta"],
                                                       o["nearestHigher"], o["isSeed"], o["clusterIndex"])
        if rc:
            raise ValueError(f"{self.prefix}clue2d failed with code {rc}")
        return o

    def clue3d(self, ev: dict, params: Clue3DParams | None = None) -> dict:
        p = params if params is not None else clue3d_params()
        n = ev["layer"].size
        o = _out3d(n)
        rc = getattr(self.lib, self.prefix + "clue3d")(
            n, _c(ev["z"], np.float32), _c(ev["eta"], np.float32), _c(ev["phi"], np.float32),
            _c(ev["r_over_absz"], np.float32), _c(ev["radius"], np.float32), _c(ev["energy"], np.float32),
            _c(ev["layer"], np.int32), _c(ev["isSilicon"], np.int32), C.byref(p),
            o["rho"], o["delta_f"], o["delta_l"], o["nh_layer"], o["nh_idx"], o["isSeed"], o["tracksterIndex"])
        if rc:
            raise ValueError(f"{self.prefix}clue3d failed with code {rc}")
        return o


class Oracle(_Base):
    prefix = "oracle_"

    def __init__(self, path: Path = LIB):
        super().__init__(path)
        for name, res, args in (("f2i", C.c_int, [C.c_float]), ("xbin", C.c_int, [C.c_float]),
                                ("ybin", C.c_int, [C.c_float]), ("reduce_range", C.c_float, [C.c_float]),
                                ("eta_bin", C.c_int, [C.c_float]), ("phi_bin", C.c_int, [C.c_float])):
            f = getattr(self.lib, "oracle_" + name)
            f.restype, f.argtypes = res, args

    def xbin(self, v):
        return self.lib.oracle_xbin(float(v))

    def ybin(self, v):
        return self.lib.oracle_ybin(float(v))

    def reduce_range(self, v):
        return np.float32(self.lib.oracle_reduce_range(float(v)))

    def eta_bin(self, v):
        return self.lib.oracle_eta_bin(float(v))

    def phi_bin(self, v):
        return self.lib.oracle_phi_bin(float(v))

    def clue2d_batch(self, packed: dict, offs, dc=3.0, rhoc=10.0, odf=2.0) -> dict:
        n = packed["x"].size
        out = _out2d(n)
        for e in range(len(offs) - 1):
            a, b = int(offs[e]), int(offs[e + 1])
            o = self.clue2d(packed["x"][a:b], packed["y"][a:b], packed["layer"][a:b], packed["weight"][a:b],
                            dc, rhoc, odf)
            for k in out:
                out[k][a:b] = o[k]
        return out

    def clue3d_batch(self, packed: dict, offs, params: Clue3DParams | None = None) -> dict:
        n = packed["layer"].size
        out = _out3d(n)
        for e in range(len(offs) - 1):
            a, b = int(offs[e]), int(offs[e + 1])
            o = self.clue3d({k: v[a:b] for k, v in packed.items()}, params)
            for k in out:
                out[k][a:b] = o[k]
        return out


class Reference(_Base):
    prefix = "ref_"

    def __init__(self, path: Path = REF_LIB):
        super().__init__(path)
        f = self.lib.ref_clue2d_batch
        f.restype = C.c_int
        f.argtypes = [C.c_int32, I64P, F32P, F32P, I32P, F32P, C.c_float, C.c_float, C.c_float,
                      F32P, F32P, I32P, I32P, I32P, C.c_int32]

    def binning(self, v) -> dict:
        """The reference's own getXBin/getYBin/etaBin/phiBin/reduceRange on each value."""
        v = _c(v, np.float32)
        n = v.size
        o = {k: np.zeros(n, np.int32) for k in ("xbin", "ybin", "eta_bin", "phi_bin")}
        o["reduce"] = np.zeros(n, np.float32)
        f = self.lib.ref_binning
        f.restype = C.c_int
        f.argtypes = [C.c_int32, F32P, I32P, I32P, I32P, I32P, F32P]
        f(n, v, o["xbin"], o["ybin"], o["eta_bin"], o["phi_bin"], o["reduce"])
        return o

    def clue2d_batch(self, packed: dict, offs, dc=3.0, rhoc=10.0, odf=2.0, threads=1) -> dict:
        n = packed["x"].size
        o = _out2d(n)
        rc = self.lib.ref_clue2d_batch(len(offs) - 1, _c(offs, np.int64), _c(packed["x"], np.float32),
                                       _c(packed["y"], np.float32), _c(packed["layer"], np.int32),
                                       _c(packed["weight"], np.float32), dc, rhoc, odf, o["rho"], o["delta"],
                                       o["nearestHigher"], o["isSeed"], o["clusterIndex"], int(threads))
        if rc:
            raise ValueError(f"ref_clue2d_batch failed with code {rc}")
        return o


def available_reference() -> bool:
    return REF_LIB.exists()


def build(with_reference: bool | None = None) -> None:
    """Compile the checkers (and the reference kernels when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    ref_root = Path(os.environ.get("CLUE_REFERENCE_SERIAL", "/root/reference/src/serial"))
    if with_reference is None:
        with_reference = ref_root.exists()
    if with_reference:
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", "ref-native", f"REF={ref_root}"], check=True)
