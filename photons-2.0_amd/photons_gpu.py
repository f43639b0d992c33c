"""Python front end of libphotons_fmm.so (ctypes over the C ABI in include/photons_fmm.h).

This mirrors the native handle API for tests and benchmarks.  It never computes
anything itself: without the built library, or without a GPU, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphotons_fmm.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "photons_fmm.h")

FP64, FP32_MIXED = 0, 1
PHASES = ["upload", "tree", "upward", "walk", "m2l", "l2l", "p2p", "download"]


class Params(C.Structure):
    _fields_ = [("boxsize", C.c_double), ("mass", C.c_double), ("soft", C.c_double), ("rs", C.c_double),
                ("rc", C.c_double), ("theta", C.c_double), ("maxleaf", C.c_int32), ("order", C.c_int32),
                ("periodic", C.c_int32), ("precision", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nleaf", C.c_int32), ("nnode", C.c_int32), ("depth", C.c_int32),
                ("pad", C.c_int32), ("p2p_pairs", C.c_int64), ("p2p_local_pairs", C.c_int64),
                ("m2l_pairs", C.c_int64), ("m2l_local_pairs", C.c_int64), ("p2p_interactions", C.c_int64),
                ("p2p_image_interactions", C.c_int64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


class PhotonsError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PhotonsError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        vp, dp, ip, bp, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int8), \
            C.POINTER(C.c_int64)
        L.photons_last_error.restype = C.c_char_p
        L.photons_create.argtypes = [C.POINTER(vp), C.POINTER(Params), C.c_int, vp]
        L.photons_destroy.argtypes = [vp]
        L.photons_set_params.argtypes = [vp, C.POINTER(Params)]
        L.photons_set_timing.argtypes = [vp, C.c_int]
        L.photons_get_phase_ms.argtypes = [vp, dp]
        L.photons_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.photons_force_host.argtypes = [vp, C.c_int64, dp, dp]
        L.photons_force_device.argtypes = [vp, C.c_int64, vp, vp]
        L.photons_sync.argtypes = [vp]
        L.photons_prepare_host.argtypes = [vp, C.c_int64, dp]
        L.photons_evaluate_host.argtypes = [vp, dp]
        L.photons_tree_perm.argtypes = [vp, ip]
        L.photons_tree_cells.argtypes = [vp, ip, ip, ip, dp, dp]
        L.photons_p2p_list.argtypes = [vp, C.c_int64, ip, ip, bp, i64p]
        L.photons_m2l_list.argtypes = [vp, C.c_int64, ip, ip, bp, i64p]
        L.photons_expansions.argtypes = [vp, dp, dp]
        for name in ("photons_p2m_batch",):
            getattr(L, name).argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, C.c_double, vp]
        L.photons_m2m_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.photons_l2l_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.photons_m2l_batch.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, vp, vp]
        L.photons_l2p_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def declared_symbols(header: str = HEADER):
    """Function names declared in include/photons_fmm.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "return", "sizeof")))


def _check(rc, what):
    if rc != 0:
        raise PhotonsError(f"{what}: {lib().photons_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_params(prm, precision=FP64) -> Params:
    d = prm if isinstance(prm, dict) else prm.as_dict()
    return Params(d["boxsize"], d["mass"], d["soft"], d["rs"], d["rc"], d["theta"], int(d["maxleaf"]),
                  int(d["order"]), int(d.get("periodic", 1)), int(precision))


class Context:
    """One GPU force context (device-resident particles, tree and lists)."""

    def __init__(self, prm, precision=FP64, device=0, stream=None):
        self.h = C.c_void_p()
        self.params = make_params(prm, precision)
        _check(lib().photons_create(C.byref(self.h), C.byref(self.params), device,
                                    C.c_void_p(stream) if stream else None), "photons_create")

    def close(self):
        if self.h:
            lib().photons_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, prm, precision=None):
        p = make_params(prm, self.params.precision if precision is None else precision)
        _check(lib().photons_set_params(self.h, C.byref(p)), "photons_set_params")
        self.params = p

    def set_timing(self, on=True):
        lib().photons_set_timing(self.h, 1 if on else 0)

    def phase_ms(self):
        ms = np.zeros(len(PHASES))
        lib().photons_get_phase_ms(self.h, _dp(ms))
        return dict(zip(PHASES, ms.tolist()))

    def force(self, pos):
        """host positions [n][3] -> short-range acceleration [n][3] (caller order)"""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        acc = np.zeros_like(pos)
        _check(lib().photons_force_host(self.h, pos.shape[0], _dp(pos), _dp(acc)), "photons_force_host")
        return acc

    def force_device(self, n, d_pos, d_acc):
        """device pointers (ints) of [n][3] doubles; queued on the context stream"""
        _check(lib().photons_force_device(self.h, int(n), C.c_void_p(d_pos), C.c_void_p(d_acc)),
               "photons_force_device")

    def sync(self):
        _check(lib().photons_sync(self.h), "photons_sync")

    def prepare(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        self._n = pos.shape[0]
        _check(lib().photons_prepare_host(self.h, pos.shape[0], _dp(pos)), "photons_prepare_host")

    def evaluate_tree_order(self):
        acc = np.zeros((self._n, 3))
        _check(lib().photons_evaluate_host(self.h, _dp(acc)), "photons_evaluate_host")
        return acc

    def stats(self):
        st = Stats()
        _check(lib().photons_get_stats(self.h, C.byref(st)), "photons_get_stats")
        return st.as_dict()

    def perm(self, n):
        p = np.zeros(n, np.int32)
        _check(lib().photons_tree_perm(self.h, _ip(p)), "photons_tree_perm")
        return p

    def cells(self):
        st = self.stats()
        nc = st["nleaf"] + st["nnode"]
        out = dict(npart=np.zeros(nc, np.int32), ipart=np.zeros(nc, np.int32), son=np.zeros((nc, 2), np.int32),
                   center=np.zeros((nc, 3)), width=np.zeros((nc, 3)), nleaf=st["nleaf"], nnode=st["nnode"])
        _check(lib().photons_tree_cells(self.h, _ip(out["npart"]), _ip(out["ipart"]), _ip(out["son"]),
                                        _dp(out["center"]), _dp(out["width"])), "photons_tree_cells")
        return out

    def _list(self, fn):
        cnt = C.c_int64()
        _check(fn(self.h, 0, None, None, None, C.byref(cnt)), "list size")
        n = cnt.value
        t, s, sh = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int8)
        _check(fn(self.h, n, _ip(t), _ip(s), sh.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(cnt)), "list")
        return t, s, sh

    def p2p_list(self):
        return self._list(lib().photons_p2p_list)

    def m2l_list(self):
        return self._list(lib().photons_m2l_list)

    def expansions(self):
        st = self.stats()
        nc = st["nleaf"] + st["nnode"]
        nt = nterms(self.params.order)
        M, L = np.zeros((nc, nt)), np.zeros((nc, nt))
        _check(lib().photons_expansions(self.h, _dp(M), _dp(L)), "photons_expansions")
        return M, L


def nterms(order):
    return (order + 1) * (order + 2) * (order + 3) // 6
