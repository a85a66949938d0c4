"""Restart and output through the C-ABI (src/output.f90, src/initial.f90:
151-166, src/diagnostic.f90), against the restart_* fixtures, in which the
reference's own statements ran the output phase every step (shots 1) and a
second run restarted from the var.dat the first one wrote.

  * nv_data_restore of the reference's var.dat, then the restarted steps:
    the reference's restarted run (links and counts exact, positions to
    1e-10 of the box, as the full-run parity tests);
  * nv_diagnostics after the same steps: the per-point curvature and
    stretching of var.dat and the ts.log quantities (1e-12 relative; the
    FP64 sums run in another order), and the ts.log line's text;
  * nv_data_dump, then nv_data_restore in a fresh context and more steps:
    bit for bit the run that was not interrupted (FP64; the direct sum's
    chunks and the octree are deterministic);
  * nv_tree_print against the tree_mesh001.log the reference's build_tree
    wrote (the tree_mesh fixture, tree_print on): every leaf's box and point
    count, value for value (the reference's list-directed text wraps its
    records; the values are compared, as a reader of the file reads them).
"""
import dataclasses

import numpy as np
import pytest

from golden_util import Fixture, restart_names

pytestmark = pytest.mark.gpu

COMPARE = ("x", "u", "u1", "u2", "u_sup", "u_mf", "infront", "behind", "closest", "closestd")


def solver(p, exact_order=False):
    from nuvort_amd.engine import FilamentSolver
    return FilamentSolver(p, check=False, exact_order=exact_order)


@pytest.mark.parametrize("name", restart_names())
def test_restore_reference_var_dat(name, tmp_path):
    fx = Fixture(name)
    path = tmp_path / "var.dat"
    path.write_bytes(fx.file("var.dat"))
    s = solver(fx.params)
    s.data_restore(str(path))
    got = s.get_state()
    ref0 = fx.state("final")
    assert got.pcount == ref0.pcount and got.itime == ref0.itime and got.t == ref0.t
    s.step(int(fx.z["cont_steps"]))
    got = s.get_state()
    s.close()
    ref = fx.state("cont")
    assert got.pcount == ref.pcount and got.itime == ref.itime
    assert np.array_equal(got.infront, ref.infront) and np.array_equal(got.behind, ref.behind)
    assert got.recon_count == ref.recon_count
    act = ref.infront != 0
    assert np.abs(got.x[act] - ref.x[act]).max() <= 1e-10 * fx.params.box_size[0]


@pytest.mark.parametrize("name", restart_names())
def test_diagnostics_and_output_files(name, tmp_path):
    fx = Fixture(name)
    s = solver(fx.params)
    s.set_state(fx.state("in"))
    s.step(fx.nsteps)
    d = s.diagnostics()
    got = s.get_state(diagnostics=True)
    # the reference's values after the same steps, from its var.dat and diag.log
    from nuvort_amd import lib as nvlib
    L = nvlib.load()
    import ctypes as C
    from nuvort_amd.engine import FilamentSolver
    from nuvort_amd.state import FilamentState
    path = tmp_path / "ref.dat"
    path.write_bytes(fx.file("var.dat"))
    ref = FilamentState.empty(got.pcount)
    curv, stretch = np.zeros(ref.pcount), np.zeros(ref.pcount)
    v = FilamentSolver._view(ref, ref.pcount)
    v.curv = curv.ctypes.data_as(nvlib._dp)
    v.stretch = stretch.ctypes.data_as(nvlib._dp)
    assert L.nv_var_read(str(path).encode(), C.byref(v)) == 0
    scale = lambda a: max(np.abs(a).max(), 1e-300)  # noqa: E731
    assert np.abs(got.extra["curv"] - curv).max() <= 1e-12 * scale(curv)
    assert np.abs(got.extra["stretch"] - stretch).max() <= 1e-12 * scale(stretch)
    last = [float(x) for x in fx.file("diag.log").decode().split("\n")[-2].split()]
    keys = ("maxu", "maxdu", "mean_u_sup", "mean_u_mf", "kappa_bar", "kappa_min", "kappa_max", "total_length",
            "avg_sep")
    for k, want in zip(keys, last):
        assert abs(d[k] - want) <= 1e-12 * max(abs(want), 1e-300), (k, d[k], want)
    assert d["tree_eval"] == int(last[9]) or fx.params.tree_theta == 0
    assert d["active"] == int(np.count_nonzero(ref.infront))
    # the ts.log line and the var%04d.log stream, written by the library
    ts = tmp_path / "ts.log"
    s.print_info(1, str(ts))
    ref_line = fx.file("ts.log").decode().splitlines()[-1]
    got_line = ts.read_text().splitlines()[-1]
    print(ref_line)
    print(got_line)
    assert [float(x) for x in got_line.split()] == pytest.approx([float(x) for x in ref_line.split()], rel=0, abs=1e-5)
    vl = tmp_path / "var.log"
    s.printf(str(vl))
    s.close()
    last_log = [f for f in fx.files() if f.startswith("var0")][-1]
    a, b = np.frombuffer(vl.read_bytes()[12:], np.uint8), np.frombuffer(fx.file(last_log)[12:], np.uint8)
    assert vl.read_bytes()[:12] == fx.file(last_log)[:12]          # t and pcount
    assert a.size == b.size


@pytest.mark.parametrize("name", restart_names())
def test_dump_restore_continue_bit_exact(name, tmp_path):
    """FP64 in the reference's summation order: the direct sum's chunking
    otherwise follows the point count at the time a step is recorded, which
    a restart from a later state sees differently (the sums then agree to
    rounding, not bit for bit)."""
    fx = Fixture(name)
    p = dataclasses.replace(fx.params, hip_precision="fp64")
    exact = p.velocity == "BS"
    whole = solver(p, exact)
    whole.set_state(fx.state("in"))
    whole.step(fx.nsteps + 2)
    want = whole.get_state()
    whole.close()
    first = solver(p, exact)
    first.set_state(fx.state("in"))
    first.step(fx.nsteps)
    first.diagnostics()
    path = tmp_path / "var.dat"
    first.data_dump(str(path))
    first.close()
    second = solver(p, exact)
    second.data_restore(str(path))
    second.step(2)
    got = second.get_state()
    second.close()
    assert got.pcount == want.pcount and got.itime == want.itime and got.t == want.t
    assert got.recon_count == want.recon_count
    for f in COMPARE:
        assert np.array_equal(getattr(got, f), getattr(want, f)), f


def tree_mesh_values(text: str) -> np.ndarray:
    """A tree_mesh file's records: x0, x1, y0, y1, z0, z1, pcount per leaf."""
    v = np.array([float(t) for t in text.split()])
    assert v.size % 7 == 0
    return v.reshape(-1, 7)


def test_tree_print_matches_reference(tmp_path):
    fx = Fixture("tree_mesh")
    s = solver(fx.params)
    s.set_state(fx.state("ghost"))
    s.construct_tree()
    out = tmp_path / "tree_mesh001.log"
    s.tree_print(str(out))
    s.close()
    ref = tree_mesh_values(fx.file("tree_mesh001.log").decode())
    got = tree_mesh_values(out.read_text())
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert int((fx.tree()["pcount"] <= 1).sum()) == ref.shape[0]
