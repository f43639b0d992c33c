"""ctypes front end of the FP64 CPU oracle (oracle/refcpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "refharness")


class OrcParams(C.Structure):
    _fields_ = [("boxsize", C.c_double), ("mass", C.c_double), ("soft", C.c_double),
                ("rs", C.c_double), ("rc", C.c_double), ("theta", C.c_double),
                ("maxleaf", C.c_int32), ("order", C.c_int32), ("periodic", C.c_int32),
                ("pad", C.c_int32)]


class OrcStats(C.Structure):
    _fields_ = [("p2p_local_pairs", C.c_int64), ("m2l_local_pairs", C.c_int64),
                ("p2p_ext_pairs", C.c_int64), ("m2l_ext_pairs", C.c_int64),
                ("m2l_calls", C.c_int64), ("p2p_ext_inter", C.c_int64),
                ("p2p_inter", C.c_int64), ("nleaf", C.c_int32), ("nnode", C.c_int32),
                ("maxdepth", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        bp = C.POINTER(C.c_int8)
        L.orc_nterms.restype = C.c_int
        L.orc_build.restype = C.c_void_p
        L.orc_build.argtypes = [C.POINTER(OrcParams), C.c_int, dp]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_tree_sizes.argtypes = [C.c_void_p, ip, ip, ip]
        L.orc_tree_export.argtypes = [C.c_void_p, ip, ip, ip, dp, dp, ip, ip, dp, dp, dp]
        L.orc_p2p_list.restype = C.c_int64
        L.orc_p2p_list.argtypes = [C.c_void_p, ip, ip, bp, C.c_int64]
        L.orc_m2l_list.restype = C.c_int64
        L.orc_m2l_list.argtypes = [C.c_void_p, ip, ip, bp, C.c_int64]
        L.orc_let_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_force.restype = C.c_int
        L.orc_force.argtypes = [C.POINTER(OrcParams), C.c_int, dp, dp, C.POINTER(OrcStats)]
        L.orc_p2m.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, dp]
        L.orc_m2m.argtypes = [C.c_int, dp, dp, dp]
        L.orc_m2l.argtypes = [C.c_int, C.c_double, dp, dp, dp]
        L.orc_l2l.argtypes = [C.c_int, dp, dp, dp]
        L.orc_l2p.argtypes = [C.c_int, dp, dp, C.c_int, dp, dp]
        L.orc_radial.argtypes = [C.c_double, C.c_double, C.c_int, dp]
        L.orc_dtensor.argtypes = [C.c_int, C.c_double, dp, dp]
        L.orc_acceptance.restype = C.c_int
        L.orc_acceptance.argtypes = [dp, dp, dp, C.c_double, C.c_double]
        L.orc_p2p_pair.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_direct.restype = C.c_int
        L.orc_kick_loga.restype = C.c_double
        L.orc_kick_loga.argtypes = [C.c_double] * 4
        L.orc_drift_loga.restype = C.c_double
        L.orc_drift_loga.argtypes = [C.c_double] * 4
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [C.POINTER(OrcParams), C.c_int, dp, dp, dp, C.c_double, C.c_double]
        L.orc_direct.argtypes = [C.POINTER(OrcParams), C.c_int, dp, C.c_int, ip, C.c_double, dp, C.c_int]
        L.orc_force_leaves.restype = C.c_int64
        L.orc_force_leaves.argtypes = [C.POINTER(OrcParams), C.c_int, dp, C.c_int, ip, dp, ip,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _bp(a):
    return a.ctypes.data_as(C.POINTER(C.c_int8))


def params(p) -> OrcParams:
    d = p if isinstance(p, dict) else p.as_dict()
    return OrcParams(d["boxsize"], d["mass"], d["soft"], d["rs"], d["rc"], d["theta"],
                     int(d["maxleaf"]), int(d["order"]), int(d.get("periodic", 1)), 0)


def nterms(order: int) -> int:
    return (order + 1) * (order + 2) * (order + 3) // 6


def force(pos, prm):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    acc = np.zeros((n, 3))
    st = OrcStats()
    rc = lib().orc_force(C.byref(params(prm)), n, _dp(pos), _dp(acc), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_force failed ({rc})")
    return acc, st.as_dict()


def force_leaves(pos, prm, leaves):
    """the force on the particles of the given target leaves only (leaf ids of the
    tree), with their P2P list lengths and the M2L entries of each leaf and its
    ancestors: returns (acc [m][3], caller index [m], p2p_len [k], m2l_len [k]); the
    particles come leaf after leaf in tree order"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    sel = np.ascontiguousarray(leaves, dtype=np.int32)
    p = params(prm)
    cap = len(sel) * p.maxleaf
    acc = np.zeros((cap, 3))
    ids = np.zeros(cap, np.int32)
    plen = np.zeros(len(sel), np.int64)
    mlen = np.zeros(len(sel), np.int64)
    m = lib().orc_force_leaves(C.byref(p), pos.shape[0], _dp(pos), len(sel), _ip(sel), _dp(acc), _ip(ids),
                               plen.ctypes.data_as(C.POINTER(C.c_int64)), mlen.ctypes.data_as(C.POINTER(C.c_int64)))
    if m < 0:
        raise RuntimeError(f"orc_force_leaves failed ({m})")
    return acc[:m], ids[:m], plen, mlen


def kick_loga(loga_i, loga_f, omega_m=0.25, omega_x=0.75):
    """int dln a / (a H) by 128-block Simpson (src/initial.c:650-666)"""
    return lib().orc_kick_loga(loga_i, loga_f, omega_m, omega_x)


def drift_loga(loga_i, loga_f, omega_m=0.25, omega_x=0.75):
    """int dln a / (a^2 H) by 128-block Simpson (src/initial.c:668-683)"""
    return lib().orc_drift_loga(loga_i, loga_f, omega_m, omega_x)


def step(pos, vel, acc, prm, dkh, dd):
    """one reference step (src/photoNs.c:208-268, PM omitted): returns new (pos, vel, acc)"""
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    acc = np.array(acc, dtype=np.float64, order="C")
    rc = lib().orc_step(C.byref(params(prm)), pos.shape[0], _dp(pos), _dp(vel), _dp(acc), dkh, dd)
    if rc != 0:
        raise RuntimeError(f"orc_step failed ({rc})")
    return pos, vel, acc


def wrap(x, boxsize):
    """src/photoNs.c:178-196: add or subtract the box length until x is in [0, L)"""
    x = np.array(x, dtype=np.float64)
    while True:
        lo, hi = x < 0.0, x >= boxsize
        if not (lo.any() or hi.any()):
            return x
        x = np.where(lo, x + boxsize, np.where(hi, x - boxsize, x))


def treepm_step(pos, vel, acc, acc_pm, prm, pm_force, dkh, dd):
    """one pass of the reference's time loop with both force parts
    (src/photoNs.c:158-254): v += acc_pm dkh; v += acc dkh; x += v dd; wrap; new acc
    (the FMM force of this module) and acc_pm (pm_force(pos), e.g. oracle.pm);
    v += acc dkh; v += acc_pm dkh.  Returns new (pos, vel, acc, acc_pm)."""
    vel = np.array(vel, dtype=np.float64)
    vel = vel + np.asarray(acc_pm) * dkh
    vel = vel + np.asarray(acc) * dkh
    pos = wrap(np.asarray(pos, dtype=np.float64) + vel * dd, params(prm).boxsize)
    acc, _ = force(pos, prm)
    acc_pm = np.asarray(pm_force(pos), dtype=np.float64)
    vel = vel + acc * dkh
    vel = vel + acc_pm * dkh
    return pos, vel, acc, acc_pm


class Tree:
    def __init__(self, pos, prm):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.prm = params(prm)
        self.h = lib().orc_build(C.byref(self.prm), self.pos.shape[0], _dp(self.pos))
        if not self.h:
            raise RuntimeError("orc_build failed")
        nl, nn, md = C.c_int32(), C.c_int32(), C.c_int32()
        lib().orc_tree_sizes(self.h, C.byref(nl), C.byref(nn), C.byref(md))
        self.nleaf, self.nnode, self.maxdepth = nl.value, nn.value, md.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def export(self):
        n, nl, nn = self.pos.shape[0], self.nleaf, self.nnode
        out = dict(perm=np.zeros(n, np.int32), leaf_ipart=np.zeros(nl, np.int32),
                   leaf_npart=np.zeros(nl, np.int32), leaf_center=np.zeros((nl, 3)),
                   leaf_width=np.zeros((nl, 3)), node_son=np.zeros((nn, 2), np.int32),
                   node_npart=np.zeros(nn, np.int32), node_split=np.zeros(nn),
                   node_center=np.zeros((nn, 3)), node_width=np.zeros((nn, 3)))
        lib().orc_tree_export(self.h, _ip(out["perm"]), _ip(out["leaf_ipart"]), _ip(out["leaf_npart"]),
                              _dp(out["leaf_center"]), _dp(out["leaf_width"]), _ip(out["node_son"]),
                              _ip(out["node_npart"]), _dp(out["node_split"]), _dp(out["node_center"]),
                              _dp(out["node_width"]))
        return out

    def _list(self, fn):
        n = fn(self.h, None, None, None, 0)
        t = np.zeros(n, np.int32)
        s = np.zeros(n, np.int32)
        sh = np.zeros(n, np.int8)
        fn(self.h, _ip(t), _ip(s), _bp(sh), n)
        return t, s, sh

    def p2p_list(self):
        return self._list(lib().orc_p2p_list)

    def m2l_list(self):
        return self._list(lib().orc_m2l_list)

    def let_size(self, sid):
        nn, nb = C.c_int64(), C.c_int64()
        lib().orc_let_size(self.h, sid, C.byref(nn), C.byref(nb))
        return nn.value, nb.value

    def fits_reference_buffers(self):
        """True when the reference's fixed-size arrays hold this tree: leaves and
        nodes within 2N/MAXLEAF (src/fmm.c:203-209) and every image tree within
        1.1x that many nodes and 1.1 N bodies (src/fmm.c:1004-1005)."""
        n = self.pos.shape[0]
        cap = int(2.0 * n / self.prm.maxleaf)
        if cap > n:
            cap = n + 1
        if self.nleaf > cap or self.nnode > cap:
            return False
        for sid in range(1, 27):
            nn, nb = self.let_size(sid)
            if nn > int(cap * 1.1) or nb > int(n * 1.1):
                return False
        return True


# ---------------------------------------------------------------- single operators

def p2m(order, pos, mass, center):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    M = np.zeros(nterms(order))
    lib().orc_p2m(order, pos.shape[0], _dp(pos), mass, _dp(np.asarray(center, np.float64)), _dp(M))
    return M


def m2m(order, d, M, tM=None):
    tM = np.zeros(nterms(order)) if tM is None else np.array(tM, np.float64)
    lib().orc_m2m(order, _dp(np.asarray(d, np.float64)), _dp(np.ascontiguousarray(M, np.float64)), _dp(tM))
    return tM


def m2l(order, rs, d, M, L=None):
    L = np.zeros(nterms(order)) if L is None else np.array(L, np.float64)
    lib().orc_m2l(order, rs, _dp(np.asarray(d, np.float64)), _dp(np.ascontiguousarray(M, np.float64)), _dp(L))
    return L


def l2l(order, d, L, tL=None):
    tL = np.zeros(nterms(order)) if tL is None else np.array(tL, np.float64)
    lib().orc_l2l(order, _dp(np.asarray(d, np.float64)), _dp(np.ascontiguousarray(L, np.float64)), _dp(tL))
    return tL


def l2p(order, L, center, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    acc = np.zeros_like(pos)
    lib().orc_l2p(order, _dp(np.ascontiguousarray(L, np.float64)), _dp(np.asarray(center, np.float64)),
                  pos.shape[0], _dp(pos), _dp(acc))
    return acc


def radial(r, rs, nmax):
    f = np.zeros(nmax + 1)
    lib().orc_radial(r, rs, nmax, _dp(f))
    return f


def dtensor(order, rs, R):
    D = np.zeros(nterms(order))
    lib().orc_dtensor(order, rs, _dp(np.asarray(R, np.float64)), _dp(D))
    return D


def acceptance(wi, wj, dist, theta, rc):
    return lib().orc_acceptance(_dp(np.asarray(wi, np.float64)), _dp(np.asarray(wj, np.float64)),
                                _dp(np.asarray(dist, np.float64)), theta, rc)


def p2p_pair(xt, xs, mass, soft, rs):
    acc = np.zeros(3)
    lib().orc_p2p_pair(_dp(np.asarray(xt, np.float64)), _dp(np.asarray(xs, np.float64)), mass, soft, rs, _dp(acc))
    return acc


def direct(pos, prm, targets, rmax, nthreads=8):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.int32)
    acc = np.zeros((len(targets), 3))
    rc = lib().orc_direct(C.byref(params(prm)), pos.shape[0], _dp(pos), len(targets), _ip(targets), rmax,
                          _dp(acc), nthreads)
    if rc != 0:
        raise RuntimeError("orc_direct refused its arguments")
    return acc


# ---------------------------------------------------------------- compiled reference (this container)

def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


def ref_force(pos, prm, workdir):
    """Run the reference FMM subset (oracle/_ref/refharness) on pos."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    d = prm if isinstance(prm, dict) else prm.as_dict()
    fin = os.path.join(workdir, "ref_in.bin")
    fout = os.path.join(workdir, "ref_out.bin")
    with open(fin, "wb") as f:
        f.write(np.array([n, d["maxleaf"], 0, 0], np.int32).tobytes())
        f.write(np.array([d["boxsize"], d["mass"], d["soft"], d["rs"], d["rc"], d["theta"]]).tobytes())
        f.write(pos.tobytes())
    subprocess.run([REF_HARNESS, "force", fin, fout], check=True)
    raw = open(fout, "rb").read()
    acc = np.frombuffer(raw[: 24 * n], dtype=np.float64).reshape(n, 3).copy()
    cnt = np.frombuffer(raw[24 * n: 24 * n + 48], dtype=np.int64)
    keys = ["idxP2P", "idxM2L", "walk_m2l_count", "p2p_count_remote", "nleaf", "nnode"]
    return acc, dict(zip(keys, [int(c) for c in cnt]))


def ref_tree_multipoles(pos, prm, workdir):
    """fmm_prepare of the reference FMM subset: the multipoles of btree[] (from
    first_node on) and of leaf[] (from first_leaf on), p = 3"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    d = prm if isinstance(prm, dict) else prm.as_dict()
    fin = os.path.join(workdir, "ref_in.bin")
    fout = os.path.join(workdir, "ref_tree.bin")
    with open(fin, "wb") as f:
        f.write(np.array([n, d["maxleaf"], 0, 0], np.int32).tobytes())
        f.write(np.array([d["boxsize"], d["mass"], d["soft"], d["rs"], d["rc"], d["theta"]]).tobytes())
        f.write(pos.tobytes())
    subprocess.run([REF_HARNESS, "tree", fin, fout], check=True)
    raw = open(fout, "rb").read()
    nleaf, nnode = np.frombuffer(raw[:16], dtype=np.int64)
    M = np.frombuffer(raw[16:], dtype=np.float64)
    node_M = M[: nnode * 20].reshape(nnode, 20).copy()
    leaf_M = M[nnode * 20:].reshape(nleaf, 20).copy()
    return node_M, leaf_M
