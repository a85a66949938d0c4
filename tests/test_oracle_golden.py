"""The oracle (oracle/nv_oracle.c) against the golden fixtures written by the
reference build: every phase of the first step and the final state after the
fixture's full run must be bit-identical."""
import numpy as np
import pytest

from golden_util import Fixture, compare_states, names
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
    for ph in ("ghost", "motion", "insert", "close", "recon", "remove"):
        if fx.has(ph):
            searched = ph in ("close", "recon", "remove")
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
