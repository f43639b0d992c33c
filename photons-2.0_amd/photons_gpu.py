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
        L.photons_set_plan_reuse.argtypes = [vp, C.c_int]
        L.photons_get_phase_ms.argtypes = [vp, dp]
        L.photons_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.photons_force_host.argtypes = [vp, C.c_int64, dp, dp]
        L.photons_force_device.argtypes = [vp, C.c_int64, vp, vp]
        L.photons_sync.argtypes = [vp]
        L.photons_force_boxes_host.argtypes = [vp, C.c_int, i64p, dp, dp, dp]
        L.photons_force_boxes_device.argtypes = [vp, C.c_int, i64p, dp, vp, vp]
        L.photons_set_p2p_stage.argtypes = [vp, C.c_int]
        L.photons_get_p2p_stage.argtypes = [vp]
        L.photons_autotune_boxes.argtypes = [vp, C.c_int, i64p, dp, vp, vp, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int), dp]
        L.photons_force_device_static.argtypes = [vp, C.c_int64, vp, vp]
        L.photons_graph_create.argtypes = [vp, C.c_int, i64p, dp, vp, vp, C.c_double, C.c_int]
        L.photons_graph_launch.argtypes = [vp]
        L.photons_graph_check.argtypes = [vp, C.POINTER(C.c_int), dp]
        L.photons_graph_destroy.argtypes = [vp]
        L.photons_prepare_host.argtypes = [vp, C.c_int64, dp]
        L.photons_evaluate_host.argtypes = [vp, dp]
        L.photons_tree_perm.argtypes = [vp, ip]
        L.photons_tree_cells.argtypes = [vp, ip, ip, ip, dp, dp]
        L.photons_p2p_list.argtypes = [vp, C.c_int64, ip, ip, bp, i64p]
        L.photons_m2l_list.argtypes = [vp, C.c_int64, ip, ip, bp, i64p]
        L.photons_expansions.argtypes = [vp, dp, dp]
        L.photons_list_lengths.argtypes = [vp, C.c_int, i64p]
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

    def set_plan_reuse(self, on=True):
        """later evaluations of the same boxes run on the first one's plan (default on)"""
        lib().photons_set_plan_reuse(self.h, 1 if on else 0)

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

    def force_boxes(self, counts, lengths, pos):
        """independent periodic boxes in one evaluation: counts[b] particles of box b
        (concatenated in pos, box 0 first) in [0, lengths[b])^3 -> accelerations"""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        if counts.shape != lengths.shape or counts.sum() != pos.shape[0]:
            raise PhotonsError("counts, lengths and positions do not match")
        acc = np.zeros_like(pos)
        _check(lib().photons_force_boxes_host(self.h, len(counts), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                              _dp(lengths), _dp(pos), _dp(acc)), "photons_force_boxes_host")
        return acc

    def force_boxes_device(self, counts, lengths, d_pos, d_acc):
        """as force_boxes with device pointers (ints); queued on the context stream"""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        self._keep = (counts, lengths)
        _check(lib().photons_force_boxes_device(self.h, len(counts), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                                _dp(lengths), C.c_void_p(d_pos), C.c_void_p(d_acc)),
               "photons_force_boxes_device")

    def graph_create(self, counts, lengths, d_pos, d_acc, margin=0.25, capture=True):
        """plan the evaluation of these boxes at device arrays d_pos -> d_acc (pointers as
        ints) and capture it into a hipGraph; d_acc holds the force of d_pos afterwards"""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        _check(lib().photons_graph_create(self.h, len(counts), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                          _dp(lengths), C.c_void_p(d_pos), C.c_void_p(d_acc), float(margin),
                                          1 if capture else 0), "photons_graph_create")

    def graph_launch(self):
        """queue one replay of the planned evaluation (returns at once)"""
        _check(lib().photons_graph_launch(self.h), "photons_graph_launch")

    def graph_check(self):
        """wait; returns (overflow bits of the last launch, its GPU ms or -1)"""
        ovf, ms = C.c_int(), C.c_double()
        _check(lib().photons_graph_check(self.h, C.byref(ovf), C.byref(ms)), "photons_graph_check")
        return ovf.value, ms.value

    def graph_destroy(self):
        _check(lib().photons_graph_destroy(self.h), "photons_graph_destroy")

    def step_device(self, pm, n, d_pos, d_vel, d_acc, d_acc_pm, dkh, dd):
        """one TreePM step of [n][3] device arrays (pointers as ints): half kicks with
        the previous acc_pm and acc, drift, wrap, new acc and acc_pm, half kicks; pm may
        be None (no long-range part, d_acc_pm unused).  Queued on the context stream."""
        L = _pm_lib()
        _check(L.photons_step_device(self.h, pm.h if pm is not None else None, int(n), C.c_void_p(d_pos),
                                     C.c_void_p(d_vel), C.c_void_p(d_acc),
                                     C.c_void_p(d_acc_pm) if d_acc_pm else None, float(dkh), float(dd)),
               "photons_step_device")

    def sync(self):
        _check(lib().photons_sync(self.h), "photons_sync")

    def set_p2p_stage(self, stage):
        _check(lib().photons_set_p2p_stage(self.h, int(stage)), "photons_set_p2p_stage")

    def p2p_stage(self):
        return lib().photons_get_p2p_stage(self.h)

    def autotune_boxes(self, counts, lengths, d_pos, d_acc, reps=1):
        """time leaf 16/32/64 x P2P staging 256/512 on this batch; keeps the fastest.
        Returns (leaf, stage, {(leaf, stage): ms})"""
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        bl, bs, ms = C.c_int(), C.c_int(), np.zeros(6)
        _check(lib().photons_autotune_boxes(self.h, len(counts), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                            _dp(lengths), C.c_void_p(d_pos), C.c_void_p(d_acc), int(reps),
                                            C.byref(bl), C.byref(bs), _dp(ms)), "photons_autotune_boxes")
        self.params.maxleaf = bl.value
        table = {(leaf, st): float(ms[2 * i + j]) for i, leaf in enumerate((16, 32, 64))
                 for j, st in enumerate((256, 512)) if ms[2 * i + j] >= 0}
        return bl.value, bs.value, table

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

    def list_lengths(self, which):
        """entries per target of the last force call: 'p2p' (per leaf) or 'm2l' (per cell)"""
        st = self.stats()
        nt = st["nleaf"] if which == "p2p" else st["nleaf"] + st["nnode"]
        out = np.zeros(nt, np.int64)
        _check(lib().photons_list_lengths(self.h, 0 if which == "p2p" else 1,
                                          out.ctypes.data_as(C.POINTER(C.c_int64))), "photons_list_lengths")
        return out

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


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_int64), ("particles_stepped", C.c_int64), ("batches", C.c_int64),
                ("queued", C.c_int64), ("mean_boxes_per_batch", C.c_double), ("span_s", C.c_double),
                ("resident_bytes", C.c_double), ("reserved_bytes", C.c_double), ("nboxes", C.c_int32),
                ("max_lanes_busy", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


def _server_lib():
    L = lib()
    if not getattr(L, "_server_types", False):
        vp, dp, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.photons_server_last_error.restype = C.c_char_p
        L.photons_server_create.argtypes = [C.POINTER(vp), C.POINTER(Params), C.c_int, C.c_int, C.c_double, C.c_int64]
        L.photons_server_destroy.argtypes = [vp]
        L.photons_server_add_box.argtypes = [vp, C.c_int64, C.c_double, dp, dp, dp, C.POINTER(C.c_int)]
        L.photons_server_add_box_pm.argtypes = [vp, C.c_int64, C.c_double, C.c_int, dp, dp, dp, dp,
                                                C.POINTER(C.c_int)]
        L.photons_server_read_pm.argtypes = [vp, C.c_int, dp]
        L.photons_server_pause.argtypes = [vp, C.c_int]
        L.photons_server_remove_box.argtypes = [vp, C.c_int]
        L.photons_server_submit.argtypes = [vp, C.c_int, C.c_double, C.c_double, i64p]
        L.photons_server_wait.argtypes = [vp, C.c_int64, dp]
        L.photons_server_drain.argtypes = [vp]
        L.photons_server_read_box.argtypes = [vp, C.c_int, dp, dp, dp]
        L.photons_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.photons_kick_factor.restype = C.c_double
        L.photons_kick_factor.argtypes = [C.c_double] * 4
        L.photons_drift_factor.restype = C.c_double
        L.photons_drift_factor.argtypes = [C.c_double] * 4
        L._server_types = True
    return L


def _scheck(rc, what):
    if rc != 0:
        raise PhotonsError(f"{what}: {_server_lib().photons_server_last_error().decode()}")


def kick_factor(loga_i, loga_f, omega_m=0.25, omega_x=0.75):
    return _server_lib().photons_kick_factor(loga_i, loga_f, omega_m, omega_x)


def drift_factor(loga_i, loga_f, omega_m=0.25, omega_x=0.75):
    return _server_lib().photons_drift_factor(loga_i, loga_f, omega_m, omega_x)


class Server:
    """Resident boxes and a continuous-batching request loop (photons_server_*)."""

    def __init__(self, prm, precision=FP32_MIXED, device=0, nlanes=2, hbm_budget=200e9,
                 max_batch_particles=32 << 20):
        L = _server_lib()
        self.h = C.c_void_p()
        self.params = make_params(prm, precision)
        _scheck(L.photons_server_create(C.byref(self.h), C.byref(self.params), device, nlanes, float(hbm_budget),
                                        int(max_batch_particles)), "photons_server_create")
        self._n = {}

    def close(self):
        if self.h:
            _server_lib().photons_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_box(self, pos, vel, boxsize, acc=None, nside=None, acc_pm=None):
        """a resident box; with nside, a PM mesh of nside^3 cells and the long-range half
        kicks in its steps (acc and acc_pm both given or both None)"""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        keep = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (acc, acc_pm)]
        a, apm = (None if x is None else _dp(x) for x in keep)
        bid = C.c_int()
        L = _server_lib()
        if nside is None:
            _scheck(L.photons_server_add_box(self.h, pos.shape[0], float(boxsize), _dp(pos), _dp(vel), a,
                                             C.byref(bid)), "photons_server_add_box")
        else:
            _scheck(L.photons_server_add_box_pm(self.h, pos.shape[0], float(boxsize), int(nside), _dp(pos), _dp(vel),
                                                a, apm, C.byref(bid)), "photons_server_add_box_pm")
        self._n[bid.value] = pos.shape[0]
        return bid.value

    def remove_box(self, bid):
        _scheck(_server_lib().photons_server_remove_box(self.h, bid), "photons_server_remove_box")
        del self._n[bid]

    def submit(self, bid, dkh, dd):
        rid = C.c_int64()
        _scheck(_server_lib().photons_server_submit(self.h, bid, float(dkh), float(dd), C.byref(rid)),
                "photons_server_submit")
        return rid.value

    def wait(self, rid):
        ms = C.c_double()
        _scheck(_server_lib().photons_server_wait(self.h, rid, C.byref(ms)), "photons_server_wait")
        return ms.value

    def drain(self):
        _scheck(_server_lib().photons_server_drain(self.h), "photons_server_drain")

    def pause(self, paused=True):
        _scheck(_server_lib().photons_server_pause(self.h, 1 if paused else 0), "photons_server_pause")

    def read_box(self, bid):
        n = self._n[bid]
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        _scheck(_server_lib().photons_server_read_box(self.h, bid, _dp(pos), _dp(vel), _dp(acc)),
                "photons_server_read_box")
        return pos, vel, acc

    def read_pm(self, bid):
        acc_pm = np.zeros((self._n[bid], 3))
        _scheck(_server_lib().photons_server_read_pm(self.h, bid, _dp(acc_pm)), "photons_server_read_pm")
        return acc_pm

    def stats(self):
        st = ServerStats()
        _scheck(_server_lib().photons_server_get_stats(self.h, C.byref(st)), "photons_server_get_stats")
        return st.as_dict()


# --------------------------------------------------------------------------- Gadget-2 files (host only)

class SnapshotInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("npart", C.c_int32 * 6), ("npart_total", C.c_int32 * 6), ("num_files", C.c_int32),
                ("pad", C.c_int32), ("mass", C.c_double * 6), ("time", C.c_double), ("redshift", C.c_double),
                ("boxsize", C.c_double), ("omega_m", C.c_double), ("omega_l", C.c_double), ("hubble", C.c_double)]

    def as_dict(self):
        d = {}
        for f in self._fields_:
            if f[0] == "pad":
                continue
            v = getattr(self, f[0])
            d[f[0]] = list(v) if hasattr(v, "__len__") else v
        return d


def _snap_lib():
    L = lib()
    if not getattr(L, "_snap_types", False):
        L.photons_snapshot_last_error.restype = C.c_char_p
        L.photons_snapshot_header.argtypes = [C.c_char_p, C.POINTER(SnapshotInfo)]
        L.photons_snapshot_read.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
        L.photons_snapshot_write.argtypes = [C.c_char_p, C.POINTER(SnapshotInfo), C.c_int64, C.c_int64,
                                             C.c_void_p, C.c_void_p]
        L._snap_types = True
    return L


def _snap_check(rc, what):
    if rc != 0:
        raise PhotonsError(f"{what}: {_snap_lib().photons_snapshot_last_error().decode()}")


def snapshot_header(path):
    """header of a Gadget-2 file as a dict (src/snapshot.c:64-120)"""
    info = SnapshotInfo()
    _snap_check(_snap_lib().photons_snapshot_header(str(path).encode(), C.byref(info)), "photons_snapshot_header")
    return info.as_dict()


def snapshot_read(path, start=0, count=None):
    """positions and velocities [count][3] (velocities times a^1.5, src/snapshot.c:211-293)"""
    if count is None:
        count = snapshot_header(path)["n"] - start
    pos, vel = np.zeros((count, 3)), np.zeros((count, 3))
    _snap_check(_snap_lib().photons_snapshot_read(str(path).encode(), int(start), int(count), pos.ctypes.data,
                                                  vel.ctypes.data), "photons_snapshot_read")
    return pos, vel


def snapshot_write(path, pos, vel, boxsize, redshift, mass, omega_m=0.25, omega_l=0.75, hubble=0.7, n_total=None):
    """one file of type-1 particles (src/snapshot.c:397-503)"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    info = SnapshotInfo()
    info.boxsize, info.redshift, info.omega_m, info.omega_l, info.hubble = boxsize, redshift, omega_m, omega_l, hubble
    info.mass[1] = mass
    n = pos.shape[0]
    _snap_check(_snap_lib().photons_snapshot_write(str(path).encode(), C.byref(info), n,
                                                   n if n_total is None else int(n_total), pos.ctypes.data,
                                                   vel.ctypes.data), "photons_snapshot_write")


# --------------------------------------------------------------------------- PM long-range force

def _pm_lib():
    L = lib()
    if not getattr(L, "_pm_types", False):
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.photons_pm_last_error.restype = C.c_char_p
        L.photons_pm_create.argtypes = [C.POINTER(vp), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, vp]
        L.photons_pm_destroy.argtypes = [vp]
        L.photons_pm_force_device.argtypes = [vp, C.c_int64, vp, vp]
        L.photons_pm_force_host.argtypes = [vp, C.c_int64, dp, dp]
        L.photons_pm_sync.argtypes = [vp]
        L.photons_step_device.argtypes = [vp, vp, C.c_int64, vp, vp, vp, vp, C.c_double, C.c_double]
        L._pm_types = True
    return L


class PM:
    """PM long-range force on an nside^3 mesh (photons_pm_*): acc_pm without G."""

    def __init__(self, nside, boxsize, rs, mass, device=0, stream=None):
        L = _pm_lib()
        self.h = C.c_void_p()
        rc = L.photons_pm_create(C.byref(self.h), int(nside), float(boxsize), float(rs), float(mass), device,
                                 C.c_void_p(stream) if stream else None)
        if rc != 0:
            raise PhotonsError(f"photons_pm_create: {L.photons_pm_last_error().decode()}")

    def close(self):
        if self.h:
            _pm_lib().photons_pm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def force(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        acc = np.zeros_like(pos)
        if _pm_lib().photons_pm_force_host(self.h, pos.shape[0], _dp(pos), _dp(acc)) != 0:
            raise PhotonsError(f"photons_pm_force_host: {_pm_lib().photons_pm_last_error().decode()}")
        return acc

    def force_device(self, n, d_pos, d_acc):
        if _pm_lib().photons_pm_force_device(self.h, int(n), C.c_void_p(d_pos), C.c_void_p(d_acc)) != 0:
            raise PhotonsError(f"photons_pm_force_device: {_pm_lib().photons_pm_last_error().decode()}")

    def sync(self):
        if _pm_lib().photons_pm_sync(self.h) != 0:
            raise PhotonsError(f"photons_pm_sync: {_pm_lib().photons_pm_last_error().decode()}")
