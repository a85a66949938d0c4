"""Test infrastructure only: state files of the reference build's driver
(oracle/ref_harness.f90) and a runner for it.

The file layout is the one ref_harness.f90 documents: int32 pcount, int32
itime, int32 recon_count, real8 t, then every per-point field as a
contiguous array (vectors component by component).
"""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np

from nuvort_amd.state import FilamentState

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "nuvort_ref")


def write_state(path: str, st: FilamentState) -> None:
    n = st.pcount
    with open(path, "wb") as fh:
        fh.write(struct.pack("<iiid", n, st.itime, st.recon_count, st.t))
        for name, arr in st.arrays():
            if arr.ndim == 2:
                for k in range(3):
                    fh.write(np.ascontiguousarray(arr[:, k]).astype(arr.dtype).tobytes())
            else:
                fh.write(np.ascontiguousarray(arr).tobytes())


def read_state(path: str) -> FilamentState:
    with open(path, "rb") as fh:
        data = fh.read()
    n, itime, rc, t = struct.unpack_from("<iiid", data, 0)
    off = 20
    st = FilamentState.empty(n)
    st.itime, st.recon_count, st.t = itime, rc, t
    for name, arr in st.arrays():
        dt = arr.dtype
        if arr.ndim == 2:
            for k in range(3):
                arr[:, k] = np.frombuffer(data, dt, n, off)
                off += n * dt.itemsize
        else:
            arr[:] = np.frombuffer(data, dt, n, off)
            off += n * dt.itemsize
    assert off == len(data), (off, len(data))
    return st


TREE_REC = np.dtype([("pos", "<f8", 3), ("width", "<f8"), ("cent", "<f8", 3), ("circ", "<f8", 3),
                     ("pcount", "<i4"), ("leaf_infront", "<i4")])


def read_tree(path: str) -> np.ndarray:
    return np.fromfile(path, TREE_REC)


def run_reference(workdir: str, st: FilamentState, run_text: str, nsteps: int, dump_phases: bool = True,
                  threads: int = 1, files: dict | None = None, expect_fatal: bool = False, output: bool = False,
                  state_path: str | None = None) -> dict:
    """Run the reference build on ``st`` for ``nsteps`` steps in ``workdir``.
    ``files`` maps file names (relative to ``workdir``) to text written there
    first, such as a noise table.  Returns the dumped states (ghost, tree,
    motion, insert, close, recon, remove for the first step; final after the
    last; wall after wall_recon), the reconnection log lines and the timing
    record.  With
    ``expect_fatal`` the run is expected to stop in fatal_error: its exit
    status and output are returned instead.  With ``output`` every step ends
    with the reference's output phase (diagnostics; var.dat, var%04d.log and
    the ts.log line every shots steps), whose files are returned as bytes
    under "files".  The tree_mesh%03d.log leaf lists build_tree writes with
    tree_print on (src/tree.f90:96-102) are returned there in every run.
    ``state_path`` (a var.dat) replaces ``st`` as the input."""
    os.makedirs(os.path.join(workdir, "data"), exist_ok=True)
    for name, text in (files or {}).items():
        with open(os.path.join(workdir, name), "w") as fh:
            fh.write(text)
    with open(os.path.join(workdir, "run.in"), "w") as fh:
        fh.write(run_text)
    log = os.path.join(workdir, "data", "recon_location.log")
    if os.path.exists(log):
        os.remove(log)
    src = "state_in.bin"
    if state_path is not None:
        src = "restart_in.dat"
        with open(state_path, "rb") as fi, open(os.path.join(workdir, src), "wb") as fo:
            fo.write(fi.read())
    else:
        write_state(os.path.join(workdir, src), st)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([REF_BIN, src, "out", str(nsteps), "1" if dump_phases else "0", "1" if output else "0"],
                       cwd=workdir, env=env, capture_output=True, text=True)
    if expect_fatal:
        # fatal_error (src/cdata.f90:245-255) prints its location and message
        # and stops the run
        return {"returncode": r.returncode, "stdout": r.stdout, "stderr": r.stderr}
    if r.returncode != 0:
        raise RuntimeError("reference build failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-2000:]))
    out = {}
    for ph in ("ghost", "motion", "insert", "close", "recon", "wall", "remove", "final"):
        p = os.path.join(workdir, "out_%s.bin" % ph)
        if os.path.exists(p):
            out[ph] = read_state(p)
    p = os.path.join(workdir, "out_tree.bin")
    if os.path.exists(p):
        out["tree"] = read_tree(p)
    # list-directed output may wrap one record over several lines: read the
    # log as a token stream of 6 values per reconnection (t, midpoint(3), angle, dl)
    vals = []
    if os.path.exists(log):
        with open(log) as fh:
            vals = [float(v) for v in fh.read().split()]
    out["recon_log"] = np.array(vals, np.float64).reshape(-1, 6)
    timing = {}
    with open(os.path.join(workdir, "out_timing.txt")) as fh:
        for line in fh:
            k, v = line.split()
            timing[k] = float(v)
    out["timing"] = timing
    files = {}
    dd = os.path.join(workdir, "data")
    for name in sorted(os.listdir(dd)):
        if name.startswith("tree_mesh") or (output and (name.startswith("var") or name in ("ts.log", "diag.log"))):
            with open(os.path.join(dd, name), "rb") as fh:
                files[name] = fh.read()
    if output or files:
        out["files"] = files
    return out
