"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker / CPU baseline.
"""
import ctypes as C
import os

import numpy as np

from viss_optimize_ws_amd import _abi as A
from viss_optimize_ws_amd.pipeline import Stages, _ptr

_LIB = None


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")


def load():
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is not built; run `make -C oracle`")
        lib = C.CDLL(p)
        A.bind(lib, "vo_", A.STAGE_FUNCS)
        A.bind(lib, "vo_", A.ORACLE_ONLY)
        lib.vo_mt64_seed.argtypes = [C.c_void_p, C.c_uint64]
        lib.vo_mt64_seed.restype = None
        lib.vo_mt64_next.argtypes = [C.c_void_p]
        lib.vo_mt64_next.restype = C.c_uint64
        lib.vo_sincos.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
        lib.vo_sincos.restype = None
        _LIB = lib
    return _LIB


class Oracle(Stages):
    def __init__(self, cfg):
        super().__init__(load(), "vo_", cfg)

    def viss_half(self, row0=0, nrows=-1, counts=False):
        n = self.nh if nrows < 0 else nrows
        out = np.zeros(n, np.complex128)
        cnt = np.zeros(n, np.int64) if counts else None
        self._check(self._lib.vo_viss_half(self._ctx, int(row0), int(nrows), _ptr(out.view(np.float64)),
                                           _ptr(cnt, C.c_int64)))
        return (out, cnt) if counts else out

    def viss_adjoint(self, res):
        res = np.ascontiguousarray(res, np.complex128)
        g = np.zeros(self.plan.npix)
        self._check(self._lib.vo_viss_adjoint(self._ctx, _ptr(res.view(np.float64)), _ptr(g)))
        return g

    def set_viss_raw(self, raw):
        raw = np.ascontiguousarray(raw, np.complex128)
        self._check(self._lib.vo_set_viss_raw(self._ctx, _ptr(raw.view(np.float64))))

    def group_segment(self):
        info = A.vw_seg_info()
        self._check(self._lib.vo_group_segment(self._ctx, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def recon_points(self, img, pts):
        pts = np.ascontiguousarray(pts, np.int64)
        cs = np.zeros(pts.size)
        ws = np.zeros(pts.size, np.uint32)
        self._check(self._lib.vo_recon_points(self._ctx, int(img), _ptr(pts, C.c_int64), pts.size, _ptr(cs),
                                              _ptr(ws, C.c_uint32)))
        return cs, ws

    def predicate(self, row0, nrows):
        out = np.zeros((nrows, self.plan.npix), np.uint8)
        self._check(self._lib.vo_get_predicate(self._ctx, int(row0), int(nrows), _ptr(out, C.c_uint8)))
        return out


def sincos_turn(x, form):
    """sin and cos of 2 pi x by the oracle's shared formulas (form 0: exact
    reduction + Taylor; form 1: 1024-entry table)."""
    x = np.ascontiguousarray(x, np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    load().vo_sincos(_ptr(x, C.c_double), C.c_int64(x.size), C.c_int32(form), _ptr(s, C.c_double),
                     _ptr(c, C.c_double))
    return s, c


class MT64:
    """mt19937_64 of the oracle (for known-answer tests)."""

    def __init__(self, seed):
        self._lib = load()
        self._buf = C.create_string_buffer(312 * 8 + 16)
        self._lib.vo_mt64_seed(self._buf, C.c_uint64(seed))

    def next(self):
        return int(self._lib.vo_mt64_next(self._buf))
