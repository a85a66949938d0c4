"""The oracle (oracle/nv_oracle.c) against the golden fixtures written by the
reference build: every phase of the first step and the final state after the
fixture's full run must be bit-identical."""
import numpy as np
import pytest

from golden_util import Fixture, compare_states, names, restart_names
from oracle import port

FIXTURES = names()


@pytest.mark.parametrize("name", FIXTURES)
def test_full_run_bit_exact(name):
    fx = Fixture(name)
    st = fx.state("in")
    nz = fx.noise()
    rc, fin, events = port.step(st, fx.params, fx.nsteps, threads=2,
                                noise=port.Noise(nz) if nz is not None else None)
    assert rc == 0
    ref = fx.state("final")
    assert compare_states(ref, fin, st.pcount) == []
    assert fin.recon_count == ref.recon_count
    assert fin.t == ref.t
    log = fx.recon_log
    assert len(events) == log.shape[0]
    for ev, row in zip(events, log):
        assert ev[3] == row[0]                       # time
        assert np.array_equal(np.array(ev[4]), row[1:4])   # midpoint of the pair
        assert ev[5] == row[4] and ev[6] == row[5]   # acos(t1.t2), l_before - l_after


@pytest.mark.parametrize("name", [n for n in FIXTURES if Fixture(n).has("ghost") or Fixture(n).has("motion")])
def test_first_step_phases_bit_exact(name):
    fx = Fixture(name)
    st = fx.state("in")
    nz = fx.noise()
    out = port.run_phases(st, fx.params, threads=2, noise=port.Noise(nz) if nz is not None else None)
    assert out["rc_motion"] == 0
    for ph in ("ghost", "motion", "insert", "close", "recon", "wall", "remove"):
        if fx.has(ph):
            searched = ph in ("close", "recon", "wall", "remove")
            assert compare_states(fx.state(ph), out[ph], st.pcount, searched=searched) == [], ph
    tree = fx.tree()
    if tree is not None:
        mine = out["tree"]
        assert len(mine) == len(tree["pcount"])
        for k in ("pos", "width", "cent", "circ", "pcount", "leaf_infront"):
            assert np.array_equal(mine[k], tree[k]), k


def test_fixtures_exercise_every_topology_path():
    """The fixture set must contain insertions into empty slots and by growth,
    removals and reconnections, under both close-point searches."""
    grew = Fixture("remesh_bs")
    assert grew.state("final").pcount > grew.state("in").pcount
    assert grew.timing["remove_count"] > 0
    assert np.count_nonzero(grew.state("in").infront == 0) >= 2
    assert Fixture("recon_bs").recon_log.shape[0] >= 1
    assert Fixture("recon_tree").recon_log.shape[0] >= 1
    assert Fixture("dense_tree_10step").recon_log.shape[0] >= 1


def test_wall_fixtures_exercise_the_wall_paths():
    """The wall fixtures pin points, cut a filament at a wall and clear a
    pinned pair (wall_recon, src/reconnection.f90:66-150), and their image
    vortices carry weight: with the walls made open, the oracle's velocities
    of the same state differ from the reference's by far more than rounding."""
    import dataclasses
    for name in ("walls_bs", "walls_tree"):
        fx = Fixture(name)
        assert fx.timing["wall_recon_count"] > 0
        ghost, motion = fx.state("ghost"), fx.state("motion")
        assert ghost.pinnedi.sum() > 0 and ghost.pinnedb.sum() > 0
        rec, wall = fx.state("recon"), fx.state("wall")
        assert np.count_nonzero(rec.infront != wall.infront) > 0
        cleared = (rec.infront != 0) & (wall.infront == 0) & (rec.pinnedb != 0)
        assert cleared.any()   # the pinned pair on the -z wall
        p = dataclasses.replace(fx.params, boundary_x="open", boundary_z="open") if name == "walls_bs" else \
            dataclasses.replace(fx.params, boundary_x="periodic", boundary_z="periodic")
        out = port.run_phases(fx.state("in"), p, threads=2)
        act = ghost.infront != 0
        du = np.abs(out["motion"].u_sup[act] - motion.u_sup[act]).max() / np.abs(motion.u_sup[act]).max()
        assert du > 1e-3, (name, du)


@pytest.mark.parametrize("name", restart_names())
def test_pinsert_total_length_matches_reference_diagnostics(name):
    """pinsert's total_length and avg_sep (src/line.f90:29-95) after every
    step, against the reference's own diag.log lines (columns 8 and 9) for the
    same steps: restart_remesh inserts into empty slots the loop has still to
    visit and leaves empty slots at its end, restart_walls skips pinned
    points.  Serial sums in the reference's row order: equal to the printed
    17 digits."""
    fx = Fixture(name)
    rows = [[float(v) for v in ln.split()] for ln in fx.file("diag.log").decode().splitlines() if ln.strip()]
    got, _ = port.step_lengths(fx.state("in"), fx.params, int(fx.z["nsteps"]), threads=1)
    assert len(got) == len(rows)
    for (tl, sep), row in zip(got, rows):
        assert tl == pytest.approx(row[7], rel=1e-15, abs=0)
        assert sep == pytest.approx(row[8], rel=1e-15, abs=0)
