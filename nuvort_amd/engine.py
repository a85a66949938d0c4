"""Host-side mirror of the reference's hot-path procedures over the HIP C-ABI.

``FilamentSolver`` keeps the filament state on the GPU and exposes one method
per reference routine, with the reference's names and call order
(src/run.f90:24-65): ``ghostp``, ``construct_tree``, ``pmotion``,
``pinsert``, ``pclose`` / ``pclose_tree``, ``precon``, ``premove``,
``enforce_boundary``, ``empty_tree``, ``nan_finder``; ``step(k)`` runs k
iterations of the main-loop body on the device.  Errors raise
``NuvortError`` where the reference calls fatal_error (src/cdata.f90:245-255).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L
from .runfile import RunParams, validate
from .state import FilamentState


# normal_velocity names (src/normal_fluid.f90:94-115) -> NV_NF_*
NF_MODELS = {"zero": 0, "xflow": 1, "noisey_xflow": 2, "shear_xflow": 3, "ychannel": 4, "channel": 5}


def read_noise(path: str) -> np.ndarray:
    """The noise table of noisey_xflow: list-directed reals, 100000 of them
    (src/normal_fluid.f90:156-162 load_noise reads exactly that many)."""
    with open(path) as fh:
        vals = np.array(fh.read().replace(",", " ").split(), dtype=np.float64)
    if vals.size < 100000:
        raise ValueError("%s holds %d values, load_noise reads 100000" % (path, vals.size))
    return vals[:100000].copy()


def nv_params(p: RunParams, exact_order: bool = False, capacity: int = 0, variant: dict | None = None) -> _L.NvParams:
    """variant: FP32C direct-sum kernel fields of nv_params (targets_per_lane,
    kahan_block, source_tile) to fix instead of the defaults / the autotuner."""
    s = _L.NvParams()
    s.delta, s.dt, s.quant_circ, s.corea = p.delta, p.dt, p.quant_circ, p.corea
    for k in range(3):
        s.box_size[k] = p.box_size[k]
        s.super_velocity[k] = p.super_velocity[k]
    for k, b in enumerate((p.boundary_x, p.boundary_y, p.boundary_z)):
        s.boundary[k] = {"open": 0, "periodic": 1, "solid": 2}.get(b, 9)
    s.velocity = {"LIA": 0, "BS": 1, "Tree": 2}[p.velocity]
    s.tree_theta = p.tree_theta
    s.tree_extra_correction = 1 if p.tree_extra_correction else 0
    s.normal_velocity = NF_MODELS.get(p.normal_velocity, 9)
    s.norm_vel_xflow = p.norm_vel_xflow
    s.alpha[0], s.alpha[1] = p.alpha
    s.normal_fluid_cutoff, s.t_zero_normal_fluid = p.normal_fluid_cutoff, p.t_zero_normal_fluid
    s.norm_shear_omega = p.norm_shear_omega
    s.recon_shots = p.recon_shots
    s.overide_timestep_check = 1 if p.overide_timestep_check else 0
    s.precision = {"fp64": 0, "fp32c": 1}[p.hip_precision]
    s.near_radius = p.hip_near_radius
    s.close_search = {"auto": 0, "brute": 1, "tree": 2}[p.hip_close_search]
    s.step_launch = {"graph": 0, "eager": 1}[p.hip_step_launch]
    s.exact_order = 1 if exact_order else 0
    s.capacity = capacity
    for k, v in (variant or {}).items():
        if k not in ("targets_per_lane", "kahan_block", "source_tile"):
            raise ValueError("not a direct-sum variant field: %s" % k)
        setattr(s, k, int(v))
    return s


class FilamentSolver:
    def __init__(self, params: RunParams, exact_order: bool = False, capacity: int = 0, check: bool = True,
                 variant: dict | None = None):
        if check:
            validate(params)
        self.params = params
        self.L = _L.load()
        self.h = C.c_void_p()
        rc = self.L.nv_create(C.byref(nv_params(params, exact_order, capacity, variant)), C.byref(self.h))
        if rc != 0:
            raise _L.NuvortError(rc, "nv_create rejected the parameters")
        self._events = []

    # ------------------------------------------------------------ plumbing
    def _call(self, name, *args):
        rc = getattr(self.L, name)(self.h, *args)
        if rc != 0:
            raise _L.NuvortError(rc, self.L.nv_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            self.L.nv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _view(st: FilamentState, cap: int | None = None) -> _L.NvState:
        v = _L.NvState()
        v.pcount = st.pcount
        v.capacity = cap if cap is not None else st.pcount
        v.itime, v.recon_count, v.t = st.itime, st.recon_count, st.t
        v.remove_count = int(st.extra.get("remove_count", 0))
        v.wall_recon_count = int(st.extra.get("wall_recon_count", 0))
        m = {"x": st.x, "u": st.u, "u1": st.u1, "u2": st.u2, "u_sup": st.u_sup, "u_mf": st.u_mf,
             "ghosti": st.ghosti, "ghostb": st.ghostb, "closestd": st.closestd}
        for k, a in m.items():
            assert a.dtype == np.float64 and a.flags.c_contiguous
            setattr(v, k, a.ctypes.data_as(_L._dp))
        for k in ("infront", "behind", "closest", "pinnedi", "pinnedb", "wpinned"):
            a = getattr(st, k)
            assert a.dtype == np.int32 and a.flags.c_contiguous
            setattr(v, k, a.ctypes.data_as(_L._ip))
        return v

    def set_state(self, st: FilamentState) -> None:
        self._keep = st
        self._call("nv_set_state", C.byref(self._view(st)))

    def get_state(self, diagnostics: bool = False) -> FilamentState:
        """The point array; with ``diagnostics`` also each point's curvature
        and stretching (as nv_diagnostics left them) in extra["curv"] and
        extra["stretch"]."""
        probe = _L.NvState()
        probe.capacity = 0
        self.L.nv_get_state(self.h, C.byref(probe))
        st = FilamentState.empty(probe.pcount)
        v = self._view(st, st.pcount)
        if diagnostics:
            st.extra["curv"] = np.zeros(st.pcount)
            st.extra["stretch"] = np.zeros(st.pcount)
            v.curv = st.extra["curv"].ctypes.data_as(_L._dp)
            v.stretch = st.extra["stretch"].ctypes.data_as(_L._dp)
        self._call("nv_get_state", C.byref(v))
        st.t, st.itime, st.recon_count = v.t, v.itime, v.recon_count
        st.extra["remove_count"] = v.remove_count
        st.extra["wall_recon_count"] = v.wall_recon_count
        return st

    def set_state_batch(self, states) -> None:
        """Upload independent tangles; each keeps its own 1-based links."""
        pcounts = [st.pcount for st in states]
        allst = FilamentState.empty(sum(pcounts))
        off = 0
        for st in states:
            n = st.pcount
            for name, arr in st.arrays():
                getattr(allst, name)[off:off + n] = arr
            off += n
        allst.t, allst.itime = states[0].t, states[0].itime
        self._keep = allst
        self._ntangles = len(states)
        pc = (C.c_int32 * len(states))(*pcounts)
        self._call("nv_set_state_batch", C.c_int32(len(states)), pc, C.byref(self._view(allst)))

    def get_state_batch(self):
        """The tangles back, one FilamentState each."""
        k = self._ntangles
        pc = (C.c_int32 * k)()
        probe = _L.NvState()
        probe.capacity = 0
        self.L.nv_get_state_batch(self.h, C.c_int32(k), pc, C.byref(probe))
        allst = FilamentState.empty(probe.pcount)
        v = self._view(allst, allst.pcount)
        self._call("nv_get_state_batch", C.c_int32(k), pc, C.byref(v))
        out, off = [], 0
        for q in range(k):
            n = pc[q]
            st = FilamentState.empty(n)
            for name, arr in st.arrays():
                arr[:] = getattr(allst, name)[off:off + n]
            st.t, st.itime = v.t, v.itime
            out.append(st)
            off += n
        return out

    def set_time(self, t: float, itime: int) -> None:
        self._call("nv_set_time", C.c_double(t), C.c_int32(itime))

    def set_noise(self, noise) -> None:
        """The noisey_xflow table (load_noise); its counter restarts at the
        first value."""
        a = np.ascontiguousarray(noise, dtype=np.float64)
        self._noise = a
        self._call("nv_set_noise", C.c_int32(a.size), a.ctypes.data_as(_L._dp))

    # ------------------------------------------------ reference procedures
    def ghostp(self):
        self._call("nv_ghostp")

    def construct_tree(self):
        self._call("nv_construct_tree")

    def pmotion(self):
        self._call("nv_pmotion")

    def velocity(self):
        self._call("nv_velocity")

    def pinsert(self):
        self._call("nv_pinsert")

    def pclose(self):
        self._call("nv_pclose")

    def pclose_tree(self):
        self._call("nv_pclose_tree")

    def precon(self):
        self._call("nv_precon")

    def wall_recon(self):
        self._call("nv_wall_recon")

    def premove(self):
        self._call("nv_premove")

    def enforce_boundary(self):
        self._call("nv_enforce_boundary")

    def empty_tree(self):
        self._call("nv_empty_tree")

    def nan_finder(self):
        self._call("nv_nan_check")

    def step(self, nsteps: int = 1):
        self._call("nv_step", C.c_int32(nsteps))

    # ---------------------------------------- diagnostics, output, restart
    def diagnostics(self) -> dict:
        """calculate_diagnostics (src/diagnostic.f90:9-18): the ts.log
        quantities; curvature and stretching are kept with the state."""
        d = _L.NvDiag()
        self._call("nv_diagnostics", C.byref(d))
        self._diag = d
        return {f: getattr(d, f) for f, _ in _L.NvDiag._fields_}

    def data_dump(self, path: str) -> None:
        """data_dump (src/output.f90:85-94): var.dat."""
        self._call("nv_data_dump", path.encode())

    def data_restore(self, path: str) -> None:
        """data_restore (src/initial.f90:151-166) from var.dat."""
        self._call("nv_data_restore", path.encode())

    def printf(self, path: str) -> None:
        """printf (src/output.f90:47-70): a var%04d.log stream file."""
        self._call("nv_printf", path.encode())

    def print_info(self, shots: int, path: str) -> None:
        """print_info's ts.log line (src/output.f90:23-37), from the last
        diagnostics()."""
        self._call("nv_print_info", C.byref(self._diag), C.c_int32(shots), path.encode())

    def tree_print(self, path: str) -> None:
        """tree_print (src/tree.f90:96-102): the current tree's leaves."""
        self._call("nv_tree_print", path.encode())

    # ------------------------------------------------------------ outputs
    def recon_events(self):
        buf = (_L.NvReconEvent * 65536)()
        n = C.c_int32(0)
        self._call("nv_get_recon_events", buf, 65536, C.byref(n))
        return [(e.itime, e.i, e.j, e.t, tuple(e.mid), e.angle, e.dl, e.tangle) for e in buf[:min(n.value, 65536)]]

    def tree_nodes(self) -> np.ndarray:
        n = C.c_int32(0)
        self._call("nv_get_tree", None, 0, C.byref(n))
        buf = (_L.NvTreeNode * max(1, n.value))()
        self._call("nv_get_tree", buf, n.value, C.byref(n))
        dt = np.dtype([("pos", "<f8", 3), ("width", "<f8"), ("cent", "<f8", 3), ("circ", "<f8", 3),
                       ("pcount", "<i4"), ("leaf_infront", "<i4")])
        return np.frombuffer(bytes(buf), dt, n.value).copy()

    def autotune(self, cache_path: str | None = None) -> dict:
        """Pick the direct-sum variant by timing it on the uploaded state
        (cached per GPU architecture and size class when a path is given)."""
        self._call("nv_autotune", cache_path.encode() if cache_path else None)
        st = self.stats()
        return {k: st[k] for k in ("targets_per_lane", "kahan_block", "source_tile", "items_per_cu", "tuned",
                                   "tune_ms")}

    def stats(self) -> dict:
        s = _L.NvStats()
        self._call("nv_get_stats", C.byref(s))
        return {f: getattr(s, f) for f, _ in _L.NvStats._fields_}
