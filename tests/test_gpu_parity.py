"""Parity of the HIP library (through its C-ABI) with the reference.

Each golden fixture (tests/golden/, written by the reference build) is
replayed phase by phase on the GPU: every phase starts from the reference's
own state before it, so integer/topology results are compared bit for bit
and velocities to the tolerances stated below.

Tolerances (relative, normalised by the field's largest magnitude):
  * FP64, reference summation order (exact_order): 1e-14 -- the terms and
    their order are the reference's; only log() in the local term
    (src/timestep.f90:79) may differ by an ulp between device and host libm.
  * FP64 chunked (the fast FP64 path): 1e-13.
  * Tree FP64: 1e-12 (internal-node moments are summed in tree order).
"""
import numpy as np
import pytest

from golden_util import Fixture, compare_states, names

pytestmark = pytest.mark.gpu

FIXTURES = names()


def solver(fx, **kw):
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(fx.params, check=False, **kw)
    nz = fx.noise()
    if nz is not None:
        s.set_noise(nz)
    return s


def rel(a, b):
    s = max(np.abs(a).max(), 1e-300)
    return float(np.abs(a - b).max() / s)


@pytest.mark.parametrize("name", FIXTURES)
def test_ghostp_bit_exact(name):
    fx = Fixture(name)
    if not fx.has("ghost"):
        pytest.skip("fixture has no ghost phase")
    s = solver(fx)
    s.set_state(fx.state("in"))
    s.ghostp()
    got = s.get_state()
    assert compare_states(fx.state("ghost"), got, fx.state("in").pcount, fields=("ghosti", "ghostb")) == []


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).params.velocity in ("BS", "LIA")])
@pytest.mark.parametrize("exact", [True, False])
def test_direct_fp64_velocity_and_ab(name, exact):
    """pmotion from the reference's ghost phase: u, u_sup, u_mf and the AB
    update (per-point Euler / AB2 / AB3 selection), direct sum or LIA."""
    fx = Fixture(name)
    if not fx.has("motion"):
        pytest.skip("fixture has no motion phase")
    src = fx.state("ghost") if fx.has("ghost") else None
    if src is None:
        pytest.skip("needs the ghost phase")
    s = solver(fx, exact_order=exact)
    s.set_state(src)
    s.pmotion()
    got = s.get_state()
    ref = fx.state("motion")
    tol = 1e-14 if exact else 1e-13
    act = ref.infront != 0
    assert rel(ref.u_sup[act], got.u_sup[act]) <= tol
    assert rel(ref.u[act], got.u[act]) <= tol
    assert rel(ref.u_mf[act], got.u_mf[act]) <= max(tol, 1e-13)
    # positions: the AB increment carries the velocity's relative error
    dx_ref = ref.x[act] - src.x[act]
    dx_got = got.x[act] - src.x[act]
    assert rel(dx_ref, dx_got) <= 1e-12
    assert np.array_equal(got.u1[act], got.u[act])


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).tree() is not None and Fixture(n).has("ghost")])
def test_tree_nodes(name):
    fx = Fixture(name)
    s = solver(fx)
    s.set_state(fx.state("ghost"))
    s.construct_tree()
    got = s.tree_nodes()
    ref = fx.tree()
    assert len(got) == len(ref["pcount"])
    # node order, boxes, counts and the leaf link: bit for bit
    for k in ("pos", "width", "pcount", "leaf_infront"):
        assert np.array_equal(got[k], ref[k]), k
    leaf = ref["pcount"] == 1
    assert np.array_equal(got["cent"][leaf], ref["cent"][leaf])
    assert np.array_equal(got["circ"][leaf], ref["circ"][leaf])
    assert rel(ref["cent"], got["cent"]) <= 1e-14
    assert rel(ref["circ"], got["circ"]) <= 1e-13


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).params.velocity == "Tree"
                                  and Fixture(n).has("ghost") and Fixture(n).has("motion")])
def test_tree_velocity(name):
    fx = Fixture(name)
    s = solver(fx)
    src = fx.state("ghost")
    s.set_state(src)
    s.construct_tree()
    s.pmotion()
    got = s.get_state()
    ref = fx.state("motion")
    act = ref.infront != 0
    assert rel(ref.u_sup[act], got.u_sup[act]) <= 1e-12
    assert rel(ref.u[act], got.u[act]) <= 1e-12
    assert rel(ref.x[act] - src.x[act], got.x[act] - src.x[act]) <= 1e-11


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("insert")])
def test_pinsert_bit_exact(name):
    fx = Fixture(name)
    s = solver(fx)
    s.set_state(fx.state("motion"))
    s.pinsert()
    got = s.get_state()
    n0 = fx.state("motion").pcount
    assert compare_states(fx.state("insert"), got, n0, searched=False) == []


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("close") and Fixture(n).has("insert")
                                  and Fixture(n).has("ghost")])
def test_close_search_bit_exact(name):
    fx = Fixture(name)
    s = solver(fx)
    if fx.params.tree_theta > 0:
        # the stale tree: built from the step's start state, searched with
        # the positions after pmotion and pinsert (src/run.f90:29-40)
        s.set_state(fx.state("ghost"))
        s.construct_tree()
        s.set_state(fx.state("insert"))
        s.pclose_tree()
    else:
        s.set_state(fx.state("insert"))
        s.pclose()
    got = s.get_state()
    ref = fx.state("close")
    n0 = fx.state("in").pcount
    assert compare_states(ref, got, n0, fields=("closest", "closestd")) == []


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("recon")])
def test_precon_bit_exact(name):
    fx = Fixture(name)
    s = solver(fx)
    st = fx.state("close")
    s.set_state(st)
    s.set_time(st.t, st.itime)
    s.precon()
    got = s.get_state()
    ref = fx.state("recon")
    assert compare_states(ref, got, fx.state("in").pcount) == []
    assert got.recon_count == ref.recon_count
    ev = s.recon_events()
    first = [r for r in fx.recon_log if r[0] == st.t]
    assert len(ev) == len(first)
    for e, row in zip(ev, first):
        assert np.array_equal(np.array(e[4]), row[1:4])
        assert e[6] == row[5]
        assert abs(e[5] - row[4]) <= 4e-16 * max(1.0, abs(row[4]))


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("wall") and Fixture(n).has("recon")])
def test_wall_recon_bit_exact(name):
    """wall_recon (src/reconnection.f90:66-150) from the reference's state
    after precon: points cut at a wall, neighbours pinned, pinned pairs
    cleared, and the counter."""
    fx = Fixture(name)
    s = solver(fx)
    s.set_state(fx.state("recon"))
    s.wall_recon()
    got = s.get_state()
    assert compare_states(fx.state("wall"), got, fx.state("in").pcount) == []


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("remove")])
def test_premove_bit_exact(name):
    fx = Fixture(name)
    s = solver(fx)
    s.set_state(fx.state("wall" if fx.has("wall") else "recon"))
    s.premove()
    got = s.get_state()
    assert compare_states(fx.state("remove"), got, fx.state("in").pcount) == []


@pytest.mark.parametrize("name", FIXTURES)
def test_full_run_matches_reference(name):
    """nsteps of the whole step on the device.  Positions drift from the
    reference only by the FP64 rounding of the velocity (chunked sums), so
    over these short horizons links and counts are identical and positions
    agree to 1e-10 of the box."""
    fx = Fixture(name)
    s = solver(fx)
    st = fx.state("in")
    s.set_state(st)
    s.step(fx.nsteps)
    got = s.get_state()
    ref = fx.state("final")
    assert got.pcount == ref.pcount
    assert np.array_equal(got.infront, ref.infront)
    assert np.array_equal(got.behind, ref.behind)
    assert got.recon_count == ref.recon_count
    assert np.array_equal(got.pinnedi, ref.pinnedi) and np.array_equal(got.pinnedb, ref.pinnedb)
    assert np.array_equal(got.wpinned, ref.wpinned)
    if "wall_recon_count" in fx.timing:
        assert got.extra["wall_recon_count"] == fx.timing["wall_recon_count"]
    act = ref.infront != 0
    assert np.abs(got.x[act] - ref.x[act]).max() <= 1e-10 * fx.params.box_size[0]
