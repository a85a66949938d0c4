"""The reference's restart and output file formats on the CPU (the library's
host-side readers and writers, no GPU work), against files the reference's
own statements wrote (restart_* fixtures: data_dump's var.dat,
src/output.f90:85-94, and printf's var%04d.log, src/output.f90:47-70)."""
import ctypes as C
import os

import numpy as np
import pytest

from golden_util import Fixture, restart_names

FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb", "infront", "behind", "closest", "closestd",
          "pinnedi", "pinnedb")


def lib():
    from nuvort_amd import lib as nvlib
    if not os.path.exists(nvlib.LIB_PATH):
        pytest.skip("library not built")
    return nvlib, nvlib.load()


def view(nvlib, st, cap, curv, stretch):
    from nuvort_amd.engine import FilamentSolver
    v = FilamentSolver._view(st, cap)
    v.curv = curv.ctypes.data_as(nvlib._dp)
    v.stretch = stretch.ctypes.data_as(nvlib._dp)
    return v


def read_var(nvlib, L, path):
    from nuvort_amd.state import FilamentState
    probe = nvlib.NvState()
    assert L.nv_var_read(path.encode(), C.byref(probe)) == 3   # capacity 0: the rows needed
    st = FilamentState.empty(probe.pcount)
    curv, stretch = np.zeros(st.pcount), np.zeros(st.pcount)
    v = view(nvlib, st, st.pcount, curv, stretch)
    assert L.nv_var_read(path.encode(), C.byref(v)) == 0
    st.t, st.itime, st.recon_count = v.t, v.itime, v.recon_count
    return st, curv, stretch


@pytest.mark.parametrize("name", restart_names())
def test_var_dat_round_trip(name, tmp_path):
    nvlib, L = lib()
    fx = Fixture(name)
    src = tmp_path / "var.dat"
    src.write_bytes(fx.file("var.dat"))
    st, curv, stretch = read_var(nvlib, L, str(src))
    ref = fx.state("final")
    assert st.pcount == ref.pcount and st.itime == ref.itime and st.t == ref.t
    assert st.recon_count == ref.recon_count
    for f in FIELDS:
        assert np.array_equal(getattr(st, f), getattr(ref, f)), f
    assert np.any(curv != 0.0) and np.any(stretch != 0.0)
    # written back: the same bytes
    out = tmp_path / "again.dat"
    v = view(nvlib, st, st.pcount, curv, stretch)
    v.t, v.itime, v.recon_count = st.t, st.itime, st.recon_count
    assert L.nv_var_write(str(out).encode(), C.byref(v)) == 0
    assert out.read_bytes() == fx.file("var.dat")
    # and printf's stream file of the same point array
    last = [f for f in fx.files() if f.startswith("var0")][-1]
    log = tmp_path / last
    assert L.nv_var_log_write(str(log).encode(), C.byref(v)) == 0
    assert log.read_bytes() == fx.file(last)


def test_var_dat_refuses_damaged_files(tmp_path):
    nvlib, L = lib()
    fx = Fixture(restart_names()[0])
    data = fx.file("var.dat")
    bad = tmp_path / "short.dat"
    bad.write_bytes(data[:len(data) // 2])
    from nuvort_amd.state import FilamentState
    big = FilamentState.empty(fx.state("final").pcount)
    c1, c2 = np.zeros(big.pcount), np.zeros(big.pcount)
    v = view(nvlib, big, big.pcount, c1, c2)
    assert L.nv_var_read(str(bad).encode(), C.byref(v)) == 1
    assert L.nv_var_read(str(tmp_path / "missing.dat").encode(), C.byref(v)) == 1
