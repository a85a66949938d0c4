"""Test infrastructure only: ctypes access to the oracle's C restatement
(libnv_oracle.so, built by ``make -C oracle port``).

Used by tests/ as the parity reference and by bench.py's CPU-baseline leg;
the product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from nuvort_amd.runfile import RunParams
from nuvort_amd.state import FilamentState

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnv_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NvoState(C.Structure):
    _fields_ = [("n", C.c_int32), ("cap", C.c_int32), ("itime", C.c_int32), ("recon_count", C.c_int32),
                ("remove_count", C.c_int32), ("wall_recon_count", C.c_int32), ("t", C.c_double),
                ("x", _dp), ("u", _dp), ("u1", _dp), ("u2", _dp), ("u_sup", _dp), ("u_mf", _dp),
                ("gi", _dp), ("gb", _dp), ("infront", _ip), ("behind", _ip), ("closest", _ip),
                ("closestd", _dp), ("pinnedi", _ip), ("pinnedb", _ip), ("wpinned", _ip)]


class NvoParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("dt", C.c_double), ("quant_circ", C.c_double), ("corea", C.c_double),
                ("box", C.c_double * 3), ("periodic", C.c_int32 * 3), ("velocity", C.c_int32),
                ("tree_theta", C.c_double), ("tree_extra_correction", C.c_int32),
                ("normal_velocity", C.c_int32), ("norm_vel_xflow", C.c_double), ("alpha", C.c_double * 2),
                ("super_velocity", C.c_double * 3), ("normal_fluid_cutoff", C.c_double),
                ("t_zero_normal_fluid", C.c_double), ("recon_shots", C.c_int32), ("threads", C.c_int32),
                ("norm_shear_omega", C.c_double), ("noise", _dp), ("noise_n", C.c_int32),
                ("noise_counter", _ip), ("solid", C.c_int32 * 3), ("pad_solid", C.c_int32)]


class NvoNode(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("width", C.c_double), ("cent", C.c_double * 3),
                ("circ", C.c_double * 3), ("pcount", C.c_int32), ("leaf_infront", C.c_int32),
                ("child", C.c_int32 * 8)]


class NvoTree(C.Structure):
    _fields_ = [("nodes", C.POINTER(NvoNode)), ("nnodes", C.c_int32), ("cap", C.c_int32)]


class NvoReconEvent(C.Structure):
    _fields_ = [("itime", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("pad", C.c_int32),
                ("t", C.c_double), ("mid", C.c_double * 3), ("angle", C.c_double), ("dl", C.c_double)]


class NvoReconLog(C.Structure):
    _fields_ = [("ev", C.POINTER(NvoReconEvent)), ("n", C.c_int32), ("cap", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE, "port"], check=True, capture_output=True)
        L = C.CDLL(LIB_PATH)
        S, P = C.POINTER(NvoState), C.POINTER(NvoParams)
        T, G = C.POINTER(NvoTree), C.POINTER(NvoReconLog)
        for name, args in {"nvo_ghostp": [S, P], "nvo_tree_build": [S, P, T], "nvo_tree_free": [T],
                           "nvo_velocity": [S, P, T, C.c_int32, C.c_int32, _dp], "nvo_pmotion": [S, P, T],
                           "nvo_ab_update": [S, P], "nvo_pinsert": [S, P], "nvo_pclose": [S, P],
                           "nvo_pclose_tree": [S, P, T], "nvo_precon": [S, P, G], "nvo_premove": [S, P],
                           "nvo_loop_killer": [S, C.c_int32], "nvo_enforce_boundary": [S, P],
                           "nvo_nan_check": [S], "nvo_step": [S, P, C.c_int32, G],
                           "nvo_wall_recon": [S, P]}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.nvo_tree_eval_count.restype = C.c_int64
        for name in ("nvo_total_length", "nvo_avg_sep"):
            getattr(L, name).argtypes = []
            getattr(L, name).restype = C.c_double
        _lib = L
    return _lib


NF_MODELS = {"zero": 0, "xflow": 1, "noisey_xflow": 2, "shear_xflow": 3, "ychannel": 4, "channel": 5}


class Noise:
    """The noisey_xflow table and its 1-based counter (normal_fluid's module
    state, src/normal_fluid.f90:29-31), shared by the calls it is passed to."""

    def __init__(self, values):
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.counter = np.ones(1, np.int32)


def params_struct(p: RunParams, threads: int = 0, noise: Noise | None = None) -> NvoParams:
    s = NvoParams()
    s.delta, s.dt, s.quant_circ, s.corea = p.delta, p.dt, p.quant_circ, p.corea
    for k in range(3):
        s.box[k] = p.box_size[k]
        s.super_velocity[k] = p.super_velocity[k]
    for k, b in enumerate((p.boundary_x, p.boundary_y, p.boundary_z)):
        s.periodic[k] = 1 if b == "periodic" else 0
        s.solid[k] = 1 if b == "solid" else 0
    s.velocity = {"LIA": 0, "BS": 1, "Tree": 2}[p.velocity]
    s.tree_theta = p.tree_theta
    s.tree_extra_correction = 1 if p.tree_extra_correction else 0
    s.normal_velocity = NF_MODELS[p.normal_velocity]
    s.norm_shear_omega = p.norm_shear_omega
    if noise is not None:
        s.noise = noise.values.ctypes.data_as(_dp)
        s.noise_n = noise.values.size
        s.noise_counter = noise.counter.ctypes.data_as(_ip)
    s.norm_vel_xflow = p.norm_vel_xflow
    s.alpha[0], s.alpha[1] = p.alpha
    s.normal_fluid_cutoff, s.t_zero_normal_fluid = p.normal_fluid_cutoff, p.t_zero_normal_fluid
    s.recon_shots = p.recon_shots
    s.threads = threads
    return s


class PortState:
    """A FilamentState copied into capacity-padded C buffers."""

    def __init__(self, st: FilamentState, cap: int | None = None):
        n = st.pcount
        cap = cap or (2 * n + 64)
        self.cap = cap
        self.arr = {}
        for name, a in st.arrays():
            b = np.zeros((cap,) + a.shape[1:], a.dtype)
            b[:n] = a
            self.arr[name] = b
        self.c = NvoState()
        self.c.n, self.c.cap, self.c.itime, self.c.recon_count = n, cap, st.itime, st.recon_count
        self.c.remove_count = 0
        self.c.t = st.t
        m = {"x": "x", "u": "u", "u1": "u1", "u2": "u2", "u_sup": "u_sup", "u_mf": "u_mf", "gi": "ghosti",
             "gb": "ghostb", "infront": "infront", "behind": "behind", "closest": "closest",
             "closestd": "closestd", "pinnedi": "pinnedi", "pinnedb": "pinnedb", "wpinned": "wpinned"}
        for cname, pname in m.items():
            a = self.arr[pname]
            ptype = _dp if a.dtype == np.float64 else _ip
            setattr(self.c, cname, a.ctypes.data_as(ptype))

    def ref(self):
        return C.byref(self.c)

    def to_state(self) -> FilamentState:
        n = self.c.n
        st = FilamentState.empty(n)
        for name, a in st.arrays():
            a[:] = self.arr[name][:n]
        st.t, st.itime, st.recon_count = self.c.t, self.c.itime, self.c.recon_count
        st.extra["remove_count"] = self.c.remove_count
        st.extra["wall_recon_count"] = self.c.wall_recon_count
        return st


class PortTree:
    def __init__(self):
        self.c = NvoTree()

    def ref(self):
        return C.byref(self.c)

    def nodes(self) -> np.ndarray:
        from .ref_io import TREE_REC
        n = self.c.nnodes
        out = np.zeros(n, TREE_REC)
        raw = np.ctypeslib.as_array(C.cast(self.c.nodes, C.POINTER(C.c_byte)), (n * C.sizeof(NvoNode),))
        rec = np.frombuffer(raw.tobytes(), np.dtype([("pos", "<f8", 3), ("width", "<f8"), ("cent", "<f8", 3),
                                                     ("circ", "<f8", 3), ("pcount", "<i4"),
                                                     ("leaf_infront", "<i4"), ("child", "<i4", 8)]))
        # reorder to depth-first pre-order (children already follow parents
        # in creation order, which is depth-first)
        for k in ("pos", "width", "cent", "circ", "pcount", "leaf_infront"):
            out[k] = rec[k]
        return out

    def raw(self) -> np.ndarray:
        n = self.c.nnodes
        raw = np.ctypeslib.as_array(C.cast(self.c.nodes, C.POINTER(C.c_byte)), (n * C.sizeof(NvoNode),))
        return np.frombuffer(raw.tobytes(), np.dtype([("pos", "<f8", 3), ("width", "<f8"), ("cent", "<f8", 3),
                                                      ("circ", "<f8", 3), ("pcount", "<i4"),
                                                      ("leaf_infront", "<i4"), ("child", "<i4", 8)]))

    def free(self):
        lib().nvo_tree_free(C.byref(self.c))


class ReconLog:
    def __init__(self, cap: int = 100000):
        self.buf = (NvoReconEvent * cap)()
        self.c = NvoReconLog()
        self.c.ev = C.cast(self.buf, C.POINTER(NvoReconEvent))
        self.c.n, self.c.cap = 0, cap

    def ref(self):
        return C.byref(self.c)

    def events(self):
        return [(e.itime, e.i, e.j, e.t, tuple(e.mid), e.angle, e.dl) for e in self.buf[:min(self.c.n, self.c.cap)]]


def step(st: FilamentState, p: RunParams, nsteps: int, threads: int = 0, noise: Noise | None = None):
    ps = PortState(st)
    log = ReconLog()
    rc = lib().nvo_step(ps.ref(), C.byref(params_struct(p, threads, noise)), nsteps, log.ref())
    return rc, ps.to_state(), log.events()


def step_lengths(st: FilamentState, p: RunParams, nsteps: int, threads: int = 0):
    """nsteps steps one at a time; pinsert's (total_length, avg_sep) after each
    (src/line.f90:29-95, the values the diagnostics print) and the final state."""
    L = lib()
    ps = PortState(st)
    log = ReconLog()
    pr = C.byref(params_struct(p, threads))
    out = []
    for _ in range(nsteps):
        rc = L.nvo_step(ps.ref(), pr, 1, log.ref())
        if rc:
            raise RuntimeError("oracle step failed with %d" % rc)
        out.append((L.nvo_total_length(), L.nvo_avg_sep()))
    return out, ps.to_state()


def ghostp(st: FilamentState, p: RunParams) -> FilamentState:
    """src/boundary.f90:10-19 ghostp (the ghosts of every point) on a copy of st."""
    ps = PortState(st)
    lib().nvo_ghostp(ps.ref(), C.byref(params_struct(p)))
    return ps.to_state()


def pinsert(st: FilamentState, p: RunParams):
    """src/line.f90:18-96 on a copy of st: (status, state, total_length, avg_sep)."""
    L = lib()
    ps = PortState(st)
    rc = L.nvo_pinsert(ps.ref(), C.byref(params_struct(p)))
    return rc, ps.to_state(), L.nvo_total_length(), L.nvo_avg_sep()


def premove(st: FilamentState, p: RunParams):
    """src/line.f90:102-130 on a copy of st."""
    ps = PortState(st)
    lib().nvo_premove(ps.ref(), C.byref(params_struct(p)))
    return ps.to_state()


def run_phases(st: FilamentState, p: RunParams, threads: int = 0, noise: Noise | None = None) -> dict:
    """One step of src/run.f90:24-65, returning the state after each phase
    (same phase names as the reference build's dumps) and the tree."""
    L = lib()
    ps = PortState(st)
    pr = C.byref(params_struct(p, threads, noise))
    out = {}
    ps.c.itime += 1
    L.nvo_ghostp(ps.ref(), pr)
    out["ghost"] = ps.to_state()
    tr = PortTree()
    use_tree = p.tree_theta > 0
    if use_tree:
        out["rc_tree"] = L.nvo_tree_build(ps.ref(), pr, tr.ref())
        out["tree"] = tr.nodes()
    out["rc_motion"] = L.nvo_pmotion(ps.ref(), pr, tr.ref() if use_tree else None)
    out["motion"] = ps.to_state()
    out["rc_insert"] = L.nvo_pinsert(ps.ref(), pr)
    out["insert"] = ps.to_state()
    log = ReconLog()
    if ps.c.itime % p.recon_shots == 0:
        if use_tree:
            L.nvo_pclose_tree(ps.ref(), pr, tr.ref())
        else:
            L.nvo_pclose(ps.ref(), pr)
        out["close"] = ps.to_state()
        L.nvo_precon(ps.ref(), pr, log.ref())
        out["recon"] = ps.to_state()
        L.nvo_wall_recon(ps.ref(), pr)
        out["wall"] = ps.to_state()
        L.nvo_premove(ps.ref(), pr)
        out["remove"] = ps.to_state()
    out["events"] = log.events()
    if use_tree:
        tr.free()
    return out
